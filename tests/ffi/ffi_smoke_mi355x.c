/* C smoke of the drop-in boundary, exercising the same contract as the
 * reference's C smoke (version, rejected render mode, rejected camera, load,
 * render, stats) plus readback. Usage: ffi_smoke_mi355x <scene.ply>. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gsplat.h"
#include "gsplat_ext.h"

static int fail(const char *what, int32_t rc) {
  fprintf(stderr, "%s failed: %s (%d): %s\n", what, gsplat_error_message(rc), rc, gsplat_last_error_message());
  return 1;
}

int main(int argc, char **argv) {
  const char *path = argc > 1 ? argv[1] : "tests/golden/minimal_ascii.ply";
  if (gsplat_version_major() != GSPLAT_API_VERSION_MAJOR_VALUE ||
      gsplat_version_minor() != GSPLAT_API_VERSION_MINOR_VALUE)
    return 2;
  GsplatConfig bad = gsplat_config_default();
  bad.mode = GSPLAT_RENDER_MODE_SORTED_ALPHA + 1;
  GsplatContext *ctx = NULL;
  if (gsplat_context_create(bad, &ctx) != GSPLAT_ERROR_INVALID_ARGUMENT || ctx != NULL) return 8;
  GsplatConfig cfg = gsplat_config_default();
  cfg.width = 64;
  cfg.height = 64;
  int32_t rc = gsplat_context_create(cfg, &ctx);
  if (rc != GSPLAT_OK || !ctx) return fail("create", rc);
  GsplatCamera cam = gsplat_camera_default();
  GsplatCamera inv = cam;
  inv.near_plane = 10.0f;
  inv.far_plane = 1.0f;
  if (gsplat_context_set_camera(ctx, inv) != GSPLAT_ERROR_INVALID_ARGUMENT) return 9;
  if ((rc = gsplat_context_set_camera(ctx, cam)) != GSPLAT_OK) return fail("set_camera", rc);
  if ((rc = gsplat_context_load_scene_path(ctx, path)) != GSPLAT_OK) return fail("load", rc);
  if ((rc = gsplat_context_render_frame(ctx)) != GSPLAT_OK) return fail("render", rc);
  GsplatStats st;
  memset(&st, 0, sizeof(st));
  if ((rc = gsplat_context_get_stats(ctx, &st)) != GSPLAT_OK) return fail("stats", rc);
  unsigned char *img = (unsigned char *)malloc(64 * 64 * 4);
  if ((rc = gsplat_context_readback_rgba8(ctx, img, 64 * 64 * 4)) != GSPLAT_OK) return fail("readback", rc);
  unsigned long sum[4] = {0, 0, 0, 0};
  for (int i = 0; i < 64 * 64; ++i)
    for (int c = 0; c < 4; ++c) sum[c] += img[4 * i + c];
  printf("ffi smoke ok\ndrawn=%u visible=%u frame_ms=%.4f means=%.4f,%.4f,%.4f,%.4f\n", st.drawn_count,
         st.visible_count, st.frame_ms, sum[0] / (64.0 * 64 * 255), sum[1] / (64.0 * 64 * 255),
         sum[2] / (64.0 * 64 * 255), sum[3] / (64.0 * 64 * 255));
  free(img);
  gsplat_context_destroy(ctx);
  return 0;
}
