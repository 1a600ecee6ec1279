// C entry points over gs::OrderReusePolicy (gsplat-rs_amd/csrc/order_reuse.h)
// for tests/test_order_reuse.py, which builds this file with g++.
#include "order_reuse.h"

extern "C" {

void *orp_new() { return new gs::OrderReusePolicy(); }
void orp_free(void *p) { delete static_cast<gs::OrderReusePolicy *>(p); }
void orp_set_interval(void *p, unsigned n) { static_cast<gs::OrderReusePolicy *>(p)->set_interval(n); }
void orp_set_diagonal(void *p, float d) { static_cast<gs::OrderReusePolicy *>(p)->set_scene_diagonal(d); }
void orp_force(void *p) { static_cast<gs::OrderReusePolicy *>(p)->force_refresh(); }
void orp_camera(void *p, const GsplatCamera *c) { static_cast<gs::OrderReusePolicy *>(p)->camera_set(*c); }
// One frame: returns 1 if it refreshes the order, 3 if that refresh was forced by the bounds, 0 to reuse.
int orp_frame(void *p, int has_order) {
  auto *o = static_cast<gs::OrderReusePolicy *>(p);
  const gs::OrderReusePolicy::Plan plan = o->plan(has_order != 0);
  o->finish(plan);
  return (plan.refresh ? 1 : 0) | (plan.fallback ? 2 : 0);
}
void orp_stats(void *p, GsplatSurfaceSortStats *out) { *out = static_cast<gs::OrderReusePolicy *>(p)->stats(); }

}  // extern "C"
