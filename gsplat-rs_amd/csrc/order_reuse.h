// order_reuse.h -- bounded-lag reuse of a depth order across the frames of a
// context (host only, no device code).
//
// The reference's SurfaceRenderSession refreshes its back-to-front order
// every `sort_interval` camera changes and otherwise draws the new camera's
// projection in the last refreshed order (render-wgpu/src/surface_session.rs:
// 556-596, 734-744, 851-860, 913-975); an order computed for an older camera
// may be presented only while it lags at most MAX_ASYNC_SORT_REVISION_LAG (2)
// camera revisions and its camera stays within the pose envelope of the
// current one: translation at most 2% of the scene's bounding-box diagonal
// (at least 1e-4) and rotation at most 0.01 rad (:17-32, 1092-1105, 1245-1262).
// A context applies both: with an interval of 1 (the default) every changed
// camera refreshes the order, exactly the plain path; a larger interval
// reuses the order while the lag and the pose envelope hold, and refreshes it
// (a fallback, flag bit 7) as soon as either would be exceeded.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "gsplat.h"

namespace gs {

class OrderReusePolicy {
 public:
  static constexpr uint64_t kMaxRevisionLag = 2;
  static constexpr float kMaxRotationRadians = 0.01f;
  static constexpr float kMaxTranslationFraction = 0.02f;

  struct Plan {
    bool refresh = true;      // sort this frame
    bool fallback = false;    // refreshed because the lag or the pose envelope ran out
  };

  // A new scene: its bounding-box diagonal sets the translation bound, and
  // the next frame refreshes.
  void set_scene_diagonal(float diagonal) {
    translation_limit_ = std::max(diagonal * kMaxTranslationFraction, 1e-4f);
    force_ = true;
  }
  // 0 is refused by the caller (INVALID_ARGUMENT); a change forces a refresh.
  void set_interval(uint32_t interval) {
    if (interval != interval_) {
      interval_ = interval;
      force_ = true;
    }
  }
  uint32_t interval() const { return interval_; }
  void force_refresh() { force_ = true; }
  // The context's camera was set: a camera that differs (bit for bit) from
  // the current one is a new revision.
  void camera_set(const GsplatCamera &c) {
    if (!has_camera_ || std::memcmp(&c, &camera_, sizeof(c)) != 0) {
      camera_ = c;
      has_camera_ = true;
      ++camera_revision_;
      dirty_ = true;
      ++changes_since_sort_;
    }
  }

  Plan plan(bool has_order) const {
    Plan p;
    if (force_ || !has_order || (dirty_ && changes_since_sort_ >= std::max<uint32_t>(interval_, 1))) return p;
    if (camera_revision_ - applied_revision_ > kMaxRevisionLag || !pose_compatible(applied_camera_, camera_)) {
      p.fallback = true;
      return p;
    }
    p.refresh = false;
    return p;
  }
  // After the frame rendered under `p`.
  void finish(const Plan &p) {
    dirty_ = false;
    force_ = false;
    refreshed_ = p.refresh;
    fallback_ = p.fallback;
    if (p.refresh) {
      changes_since_sort_ = 0;
      applied_revision_ = camera_revision_;
      applied_camera_ = camera_;
    }
  }

  // Telemetry of the last frame in the layout of GsplatSurfaceSortStats
  // (gsplat.h; flags bit 0 order refreshed, bit 1 order uploaded, bit 7
  // refresh forced by the lag or pose bounds).
  GsplatSurfaceSortStats stats() const {
    GsplatSurfaceSortStats s{};
    s.camera_revision = camera_revision_;
    s.applied_order_revision = applied_revision_;
    const uint64_t lag = camera_revision_ - applied_revision_;
    s.presented_order_revision_lag = lag > UINT32_MAX ? UINT32_MAX : (uint32_t)lag;
    s.flags = (refreshed_ ? 1u : 0u) | (refreshed_ ? 2u : 0u) | (fallback_ ? 1u << 7 : 0u);
    return s;
  }

  // async_order_pose_compatible (surface_session.rs:1245-1262), f32.
  bool pose_compatible(const GsplatCamera &a, const GsplatCamera &b) const {
    const float dx = b.position[0] - a.position[0], dy = b.position[1] - a.position[1],
                dz = b.position[2] - a.position[2];
    if (dx * dx + dy * dy + dz * dz > translation_limit_ * translation_limit_) return false;
    const float *p = a.rotation_xyzw, *q = b.rotation_xyzw;
    const float dot = std::min(std::max(std::fabs(p[0] * q[0] + p[1] * q[1] + p[2] * q[2] + p[3] * q[3]), 0.0f), 1.0f);
    return 2.0f * std::acos(dot) <= kMaxRotationRadians;
  }

 private:
  uint32_t interval_ = 1;
  float translation_limit_ = 1e-4f;
  bool force_ = true, dirty_ = true, has_camera_ = false;
  bool refreshed_ = false, fallback_ = false;
  uint32_t changes_since_sort_ = 0;
  uint64_t camera_revision_ = 0, applied_revision_ = 0;
  GsplatCamera camera_{}, applied_camera_{};
};

}  // namespace gs
