"""Batched frames (gsplat_batch_*) through the C ABI.

Every view of a batch must give exactly the frame a context holding the
same scene, camera and size renders (same visible set, permutation, tile
counts, and bit-identical fp32/RGBA8 images), whatever else the batch holds;
selected views are also checked against the oracle to the plain-path bars.
"""

import sys

import numpy as np
import pytest

import synth
from helpers import cam_dict, oracle_scene

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def single(abi, scene, cam, w, h):
    ctx = abi.Context(w, h)
    ctx.load(scene)
    ctx.set_camera(cam)
    ctx.render()
    out = dict(order=ctx.sorted_indices(scene.count), tiles=ctx.tile_counts(), img32=ctx.rgba32f(),
               img8=ctx.rgba8(), visible=ctx.stats().visible_count, pairs=ctx.stage_times().tile_pairs)
    ctx.close()
    return out


def batch_view(batch, v, count):
    return dict(order=batch.sorted_indices(v, count), tiles=batch.tile_counts(v), img32=batch.rgba32f(v),
                img8=batch.rgba8(v), visible=batch.view_stats(v).visible_count)


def assert_same(a, b, label):
    assert a["visible"] == b["visible"], f"{label}: visible {a['visible']} vs {b['visible']}"
    assert np.array_equal(a["order"], b["order"]), f"{label}: permutation differs"
    assert np.array_equal(a["tiles"], b["tiles"]), f"{label}: tile counts differ"
    assert np.array_equal(a["img32"].view(np.uint32), b["img32"].view(np.uint32)), f"{label}: fp32 image differs"
    assert np.array_equal(a["img8"], b["img8"]), f"{label}: rgba8 image differs"


def mixed_batch(abi):
    """Scenes of SH degree 0..4 and ragged sizes, one empty, at mixed resolutions."""
    specs = [(5000, 0), (12000, 3), (4097, 1), (0, 0), (9000, 2), (3000, 4), (1, 0), (20000, 3)]
    scenes = [synth.tiny_scene(n, 7000 + i, sh_degree=d) for i, (n, d) in enumerate(specs)]
    rng = np.random.default_rng(99)
    sizes = [(640, 480), (320, 240), (257, 16), (100, 75), (1280, 720), (17, 33), (64, 64), (800, 600)]
    views = []
    for v in range(12):
        s = v % len(scenes)
        w, h = sizes[(v * 3) % len(sizes)]
        pos = synth.sphere_point(rng, float(rng.uniform(2.0, 4.0)))
        views.append((s, abi.look_at(pos), w, h))
    return scenes, views


def test_batch_views_equal_single_context(abi):
    scenes, views = mixed_batch(abi)
    dev = [abi.Scene(s) for s in scenes]
    b = abi.Batch()
    say(f"[batch mixed] {len(views)} views")
    b.render([(dev[s], cam, w, h) for (s, cam, w, h) in views])
    b.wait()
    st = b.stage_times()
    assert st.tiles_y == 1
    total_pairs = 0
    for v, (s, cam, w, h) in enumerate(views):
        ref = single(abi, scenes[s], cam, w, h)
        got = batch_view(b, v, scenes[s].count)
        assert_same(got, ref, f"view {v} (scene {s}, {w}x{h})")
        total_pairs += ref["pairs"]
    assert st.tile_pairs == total_pairs
    assert st.splat_count == sum(scenes[s].count for (s, _, _, _) in views)
    b.close()


def test_batch_view_alone_equals_in_mix(abi):
    scenes, views = mixed_batch(abi)
    dev = [abi.Scene(s) for s in scenes]
    full = abi.Batch()
    full.render([(dev[s], cam, w, h) for (s, cam, w, h) in views])
    alone = abi.Batch()
    for v in (0, 5, 11):
        s, cam, w, h = views[v]
        say(f"[batch alone] view {v}")
        alone.render([(dev[s], cam, w, h)])
        assert_same(batch_view(alone, 0, scenes[s].count), batch_view(full, v, scenes[s].count), f"view {v}")
    # the same view twice in one batch gives the same frame twice
    s, cam, w, h = views[1]
    twice = abi.Batch()
    twice.render([(dev[s], cam, w, h), (dev[s], cam, w, h)])
    assert_same(batch_view(twice, 0, scenes[s].count), batch_view(twice, 1, scenes[s].count), "twice")


def test_batch_against_oracle(abi, oracle):
    scene = synth.tiny_scene(16384, 4242, sh_degree=3)
    other = synth.tiny_scene(6000, 4343, sh_degree=1)
    dev, dev_o = abi.Scene(scene), abi.Scene(other)
    cams = [abi.look_at((0.3, -0.2, -3.0)), abi.look_at((2.5, 1.0, 1.5)), abi.look_at((0.0, 3.0, 0.1))]
    views = [(dev, cams[0], 640, 480), (dev_o, cams[1], 320, 240), (dev, cams[2], 300, 200), (dev, cams[1], 64, 48)]
    b = abi.Batch()
    b.render(views)
    osc = oracle_scene(oracle, scene)
    for v in (0, 2, 3):
        _, cam, w, h = views[v]
        say(f"[batch oracle] view {v} {w}x{h}")
        cd = cam_dict(cam)
        keys_ref, idx_ref = oracle.preprocess(osc, cd)
        order_ref = oracle.sort_values_by_keys(keys_ref, idx_ref)
        assert np.array_equal(b.sorted_indices(v, scene.count), order_ref)
        proj_ref = oracle.project(osc, cd, w, h)
        assert np.array_equal(b.tile_counts(v), oracle.tile_counts(proj_ref, w, h))
        ref32, vis = oracle.render(osc, cd, w, h)
        assert vis == b.view_stats(v).visible_count
        assert np.abs(b.rgba32f(v) - ref32).max() <= 2e-3


def test_batch_pair_overflow_rerender(abi, oracle, monkeypatch):
    # Large splats binned straight into 16x16 tiles: ~5.6M pairs over the two
    # views, beyond the initial capacity of 2^20 pairs for this splat count,
    # so the frame overflows and is rendered again with grown buffers.
    monkeypatch.setenv("GSPLAT_BIN_SHIFT", "0")
    scene = synth.tiny_scene(6000, 5151, sh_degree=0, spread=0.4, log_scale=-2.2)
    dev = abi.Scene(scene)
    cam = abi.look_at((0.0, 0.0, -2.5))
    views = [(dev, cam, 1920, 1080), (dev, abi.look_at((1.0, 0.5, -2.0)), 1920, 1080)]
    b = abi.Batch()
    say("[batch overflow]")
    b.render(views)
    st = b.stage_times()
    assert st.tile_pairs > (1 << 20)
    osc = oracle_scene(oracle, scene)
    for v, (_, c, w, h) in enumerate(views):
        proj_ref = oracle.project(osc, cam_dict(c), w, h)
        assert np.array_equal(b.tile_counts(v), oracle.tile_counts(proj_ref, w, h)), f"view {v}"


def test_batch_errors(abi):
    scene = abi.Scene(synth.tiny_scene(100, 1))
    b = abi.Batch()
    with pytest.raises(abi.GsplatError) as e:
        b.render([])
    assert e.value.code == abi.INVALID_ARGUMENT
    with pytest.raises(abi.GsplatError) as e:
        b.render([(scene, abi.look_at((0, 0, -3)), 64, 64)] * 257)
    assert e.value.code == abi.INVALID_ARGUMENT
    with pytest.raises(abi.GsplatError) as e:
        b.render([(None, abi.look_at((0, 0, -3)), 64, 64)])
    assert e.value.code == abi.INVALID_ARGUMENT and "scene is null" in abi.last_error()
    bad = abi.look_at((0, 0, -3))
    bad.near_plane = 0.0
    with pytest.raises(abi.GsplatError) as e:
        b.render([(scene, bad, 64, 64)])
    assert e.value.code == abi.INVALID_ARGUMENT and "invalid camera" in abi.last_error()
    with pytest.raises(abi.GsplatError) as e:
        b.render([(scene, abi.look_at((0, 0, -3)), 0, 64)])
    assert e.value.code == abi.INVALID_ARGUMENT
    with pytest.raises(abi.GsplatError) as e:
        b.render([(scene, abi.look_at((0, 0, -3)), 20000, 64)])
    assert e.value.code == abi.UNSUPPORTED
    out = np.zeros(64 * 64 * 4, np.uint8)
    rc = abi.lib().gsplat_batch_readback_rgba8(b.h, 0, out.ctypes.data, out.size)
    assert rc == abi.SCENE_NOT_LOADED  # no frame yet
    b.render([(scene, abi.look_at((0, 0, -3)), 64, 64)])
    rc = abi.lib().gsplat_batch_readback_rgba8(b.h, 1, out.ctypes.data, out.size)
    assert rc == abi.INVALID_ARGUMENT and "out of range" in abi.last_error()
    rc = abi.lib().gsplat_batch_readback_rgba8(b.h, 0, out.ctypes.data, out.size - 1)
    assert rc == abi.INVALID_ARGUMENT
    # the scene handle may go before the batch that rendered it
    scene.close()
    b.wait()
    b.close()


def test_batch_reduce_then_scan_sorts_equal_single(abi, monkeypatch):
    # Both sorts of the batch in reduce-then-scan mode; each view is then
    # compared with a context frame under the same mode.
    monkeypatch.setenv("GSPLAT_SORT_RTS_MIN_TILES", "1")
    scenes, views = mixed_batch(abi)
    dev = [abi.Scene(s) for s in scenes]
    b = abi.Batch()
    say("[batch rts]")
    b.render([(dev[s], cam, w, h) for (s, cam, w, h) in views])
    for v, (s, cam, w, h) in enumerate(views):
        assert_same(batch_view(b, v, scenes[s].count), single(abi, scenes[s], cam, w, h), f"view {v}")


def test_autotune_keeps_images(abi):
    scenes, views = mixed_batch(abi)
    dev = [abi.Scene(s) for s in scenes]
    vs = [(dev[s], cam, w, h) for (s, cam, w, h) in views[:6]]
    b = abi.Batch()
    b.render(vs)
    before = [batch_view(b, v, scenes[views[v][0]].count) for v in range(len(vs))]
    say("[autotune batch]")
    best = b.autotune(vs, reps=2)
    assert best.bin_shift <= 3 and best.sort_rts_min_tiles in (128, 256, 1024) and best.frame_ms > 0
    assert best.raster_variant in (1, 2, 3, 4) and best.blend_min in (32, 64, 128, 256)
    t = b.tuning()
    assert (t.bin_shift, t.sort_rts_min_tiles, t.raster_variant, t.blend_min) == (
        best.bin_shift, best.sort_rts_min_tiles, best.raster_variant, best.blend_min)
    for v in range(len(vs)):
        assert_same(batch_view(b, v, scenes[views[v][0]].count), before[v], f"view {v} after autotune")
    # every setting renders the same frame; rows2_min_tiles 1 with one
    # binning layout (bin_auto off) switches variant 1 to two rows per thread
    for shift, variant, blend_min, bin_auto, rows2 in ((0, 0, 64, 1, 0), (3, 0, 256, 1, 0), (2, 2, 1, 1, 0),
                                                       (1, 2, 64, 1, 0), (2, 1, 200, 1, 0), (2, 1, 64, 0, 1),
                                                       (1, 1, 64, 0, 1), (0, 1, 64, 0, 1), (2, 3, 64, 1, 0),
                                                       (0, 3, 1, 0, 0), (2, 4, 64, 1, 0), (0, 4, 1, 0, 0),
                                                       (3, 4, 256, 1, 0)):
        t.bin_shift, t.raster_variant, t.blend_min, t.bin_auto, t.rows2_min_tiles = shift, variant, blend_min, \
            bin_auto, rows2
        b.set_tuning(t)
        assert b.tuning().rows2_min_tiles == rows2
        b.render(vs)
        for v in range(len(vs)):
            assert_same(batch_view(b, v, scenes[views[v][0]].count), before[v], f"view {v} shift {shift}")
    bad = b.tuning()
    bad.bin_shift = 4
    with pytest.raises(abi.GsplatError) as e:
        b.set_tuning(bad)
    assert e.value.code == abi.INVALID_ARGUMENT
    bad = b.tuning()
    bad.blend_min = 0
    with pytest.raises(abi.GsplatError) as e:
        b.set_tuning(bad)
    assert e.value.code == abi.INVALID_ARGUMENT
    bad = b.tuning()
    bad.raster_variant = 7
    with pytest.raises(abi.GsplatError) as e:
        b.set_tuning(bad)
    assert e.value.code == abi.INVALID_ARGUMENT


def test_context_autotune(abi):
    scene = synth.tiny_scene(20000, 3131, sh_degree=3)
    ctx = abi.Context(320, 240)
    ctx.load(scene)
    ctx.set_camera(abi.look_at((0.4, 0.2, -2.6)))
    ctx.render()
    img = ctx.rgba8()
    say("[autotune context]")
    best = ctx.autotune(reps=2)
    assert best.frame_ms > 0
    assert np.array_equal(ctx.rgba8(), img)
    ctx.close()


def test_stage_timing_switch(abi):
    # Stage times come from frames with stage events; without them only the
    # total is measured. The frame itself is the same either way.
    scene = synth.tiny_scene(20000, 4242, sh_degree=3)
    dev = abi.Scene(scene)
    views = [(dev, abi.look_at((0.3, -0.2, -2.8)), 320, 240), (dev, abi.look_at((2.0, 1.0, 1.5)), 256, 144)]
    b = abi.Batch()
    b.render(views)
    st = b.stage_times()
    base = batch_view(b, 1, scene.count)
    assert st.total_ms > 0 and (st.project_ms, st.sort_ms, st.bin_ms, st.raster_ms) == (0, 0, 0, 0)
    b.set_stage_timing(True)
    b.render(views)
    st = b.stage_times()
    assert min(st.project_ms, st.sort_ms, st.bin_ms, st.raster_ms) > 0
    assert st.project_ms + st.sort_ms + st.bin_ms + st.raster_ms <= st.total_ms * 1.0001
    assert_same(batch_view(b, 1, scene.count), base, "stage-timed frame")
    b.set_stage_timing(False)
    b.render(views)
    assert b.stage_times().raster_ms == 0
    b.close()
    # a context times its stages by default (its stats carry the phase times)
    ctx = abi.Context(320, 240)
    ctx.load(scene)
    ctx.set_camera(abi.look_at((0.3, -0.2, -2.8)))
    ctx.render()
    ctx.wait()
    s = ctx.stage_times()
    assert min(s.project_ms, s.sort_ms, s.bin_ms, s.raster_ms) > 0
    img = ctx.rgba8()
    ctx.set_stage_timing(False)
    ctx.render()
    ctx.wait()
    assert ctx.stage_times().sort_ms == 0 and ctx.stage_times().total_ms > 0
    assert np.array_equal(ctx.rgba8(), img)
    ctx.close()


def test_sort_mode_switches_keep_frames(abi):
    # One batch object alternating between a view whose depth sort runs in
    # reduce-then-scan mode (1.1M splats, 269 sort tiles) and views sorted by
    # look-back (50k, then 700k splats): every frame must equal a fresh
    # batch's frame of the same view (stale reduce-then-scan scan state must
    # never reach a later look-back sort).
    scenes = [synth.tiny_scene(n, 6100 + i, sh_degree=0, spread=0.8, log_scale=-4.0)
              for i, n in enumerate((1_100_000, 50_000, 700_000))]
    dev = [abi.Scene(s) for s in scenes]
    cams = [abi.look_at((0.3, 0.2, -3.0)), abi.look_at((2.0, 0.5, 2.0)), abi.look_at((-1.0, 1.5, -2.5))]
    sizes = [(1280, 720), (640, 480), (320, 240)]
    fresh = {}
    for s in range(3):
        f = abi.Batch()
        f.render([(dev[s], cams[s], *sizes[s])])
        fresh[s] = batch_view(f, 0, scenes[s].count)
        f.close()
    b = abi.Batch()
    for cycle in range(4):
        for s in (0, 1, 2):
            say(f"[sort mode switches] cycle {cycle} scene {s}")
            b.render([(dev[s], cams[s], *sizes[s])])
            assert_same(batch_view(b, 0, scenes[s].count), fresh[s], f"cycle {cycle} scene {s}")
    b.close()
