"""Bucket depth sort of one-view frames against the radix passes and the oracle.

One view of at most 2^19 slots sorts its depth keys by bucket: the key range
cut into 8192 fine bins, consecutive bins grouped into buckets of about 1024
keys, one stable 9-bit pass by bucket, then every bucket sorted in LDS
(sort.h bucket_sort). The order must be the radix passes' order bit for bit
(key descending, ties by splat index). A bucket of more than 4096 keys (a
fine bin of more than about 3072 equal or near-equal keys) overflows: the
frame is rendered again with the radix passes and the next frames keep them
for a while; the result must still be exact.
"""

import sys

import numpy as np
import pytest

import synth
from helpers import cam_dict, oracle_scene

pytestmark = pytest.mark.gpu

SORT_LOOKBACK, SORT_BUCKET = 0, 2  # GsplatStageTimes.depth_sort


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def frame(abi, scene, cam, w, h):
    ctx = abi.Context(w, h)
    ctx.load(scene)
    ctx.set_camera(cam)
    ctx.render()
    return ctx


def outputs(ctx, n):
    return ctx.sorted_indices(n), ctx.tile_counts(), ctx.rgba32f(), ctx.rgba8()


def both_modes(abi, monkeypatch, scene, cam, w, h):
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("GSPLAT_SORT_BUCKET", mode)
        ctx = frame(abi, scene, cam, w, h)
        out[mode] = (ctx.stage_times().depth_sort,) + outputs(ctx, scene.count)
        ctx.close()
    return out


@pytest.mark.parametrize("which", ["640x480-16k", "640x480-16k-outliers", "1080p-120k", "1080p-300k"])
def test_bucket_sort_equals_radix_passes(gpu, monkeypatch, which):
    if which.startswith("640x480-16k"):
        scene, view = synth.plain_small()
        if which.endswith("outliers"):
            # a few splats far out stretch the key range (fine bins grow coarse)
            scene.positions[:4] *= np.float32(300.0)
    elif which == "1080p-120k":
        scene, view = synth.plain_1080p(n=120_000, seed_offset=3)
    else:
        scene, view = synth.plain_1080p()
    say(f"[bucket vs radix] {which}")
    out = both_modes(gpu, monkeypatch, scene, gpu.look_at(view.position, view.target), view.width, view.height)
    assert out["1"][0] == SORT_BUCKET and out["0"][0] == SORT_LOOKBACK
    for name, a, b in zip(("order", "tile counts", "rgba32f", "rgba8"), out["1"][1:], out["0"][1:]):
        assert np.array_equal(a, b), f"{which}: {name} differ between the bucket sort and the radix passes"


def plane_scene(n_plane, n_spread, seed):
    """n_plane splats at depth exactly 4 under the default camera (identity
    rotation: the depth is the z coordinate), so their keys are equal and
    fill one bucket, plus n_spread splats over depths 1..30 that widen the
    key range."""
    s = synth.tiny_scene(n_plane + n_spread, seed, sh_degree=1, log_scale=-4.0)
    rng = np.random.default_rng(seed)
    s.positions[:, 0:2] = rng.uniform(-1.5, 1.5, (n_plane + n_spread, 2)).astype(np.float32)
    s.positions[:n_plane, 2] = np.float32(4.0)
    s.positions[n_plane:, 2] = rng.uniform(1.0, 30.0, n_spread).astype(np.float32)
    return s


def test_bucket_overflow_renders_again_exactly(gpu, oracle, monkeypatch):
    # 16,000 equal keys in one fine bin, so one bucket of more than 4096: the
    # bucket sort flags the frame, which is rendered again with the radix
    # passes.
    scene = plane_scene(16_000, 4_000, 11)
    cam = gpu.lib().gsplat_camera_default()
    w, h = 320, 240
    monkeypatch.setenv("GSPLAT_SORT_BUCKET", "1")
    say("[bucket overflow] 16000 equal keys")
    ctx = frame(gpu, scene, cam, w, h)
    assert ctx.stage_times().depth_sort == SORT_LOOKBACK, "the overflowing frame was not rendered again"
    osc = oracle_scene(oracle, scene)
    keys_ref, idx_ref = oracle.preprocess(osc, cam_dict(cam))
    assert (keys_ref == np.float32(4.0).view(np.uint32)).sum() == 16_000
    order_ref = oracle.sort_values_by_keys(keys_ref, idx_ref)
    got = outputs(ctx, scene.count)
    assert np.array_equal(got[0], order_ref), "order after the overflow differs from the oracle"
    # the next frames keep the radix passes, with the same result
    ctx.render()
    assert ctx.stage_times().depth_sort == SORT_LOOKBACK
    assert all(np.array_equal(a, b) for a, b in zip(outputs(ctx, scene.count), got))
    ctx.close()
    monkeypatch.setenv("GSPLAT_SORT_BUCKET", "0")
    ref = frame(gpu, scene, cam, w, h)
    assert all(np.array_equal(a, b) for a, b in zip(outputs(ref, scene.count), got))
    ref.close()


def test_bucket_sort_ties_below_capacity(gpu, oracle, monkeypatch):
    # 2,000 equal keys (one fine bin; with the bins before it in its bucket
    # at most 1023 + 2000 keys, below the 4096 capacity) and a spread: equal
    # keys must come out by splat index.
    scene = plane_scene(2_000, 9_000, 12)
    cam = gpu.lib().gsplat_camera_default()
    monkeypatch.setenv("GSPLAT_SORT_BUCKET", "1")
    say("[bucket ties] 2000 equal keys")
    ctx = frame(gpu, scene, cam, 320, 240)
    assert ctx.stage_times().depth_sort == SORT_BUCKET
    osc = oracle_scene(oracle, scene)
    keys_ref, idx_ref = oracle.preprocess(osc, cam_dict(cam))
    assert np.array_equal(ctx.sorted_indices(scene.count), oracle.sort_values_by_keys(keys_ref, idx_ref))
    ctx.close()


def test_bucket_sort_single_key_value_and_culled(gpu, oracle, monkeypatch):
    # every visible key equal (range 0, shift 0) with culled splats around
    # them, and a frame with every splat culled
    monkeypatch.setenv("GSPLAT_SORT_BUCKET", "1")
    cam = gpu.lib().gsplat_camera_default()
    scene = plane_scene(3_000, 0, 13)
    scene.positions[::3, 2] = np.float32(-2.0)  # a third behind the camera
    ctx = frame(gpu, scene, cam, 200, 150)
    assert ctx.stage_times().depth_sort == SORT_BUCKET
    osc = oracle_scene(oracle, scene)
    keys_ref, idx_ref = oracle.preprocess(osc, cam_dict(cam))
    assert len(idx_ref) == 2_000
    assert np.array_equal(ctx.sorted_indices(scene.count), oracle.sort_values_by_keys(keys_ref, idx_ref))
    ctx.close()
    scene.positions[:, 2] = np.float32(-2.0)
    ctx = frame(gpu, scene, cam, 200, 150)
    assert ctx.stats().visible_count == 0 and not ctx.rgba8().any()
    ctx.close()


@pytest.mark.parametrize("n,mode", [(1 << 19, SORT_BUCKET), ((1 << 19) + 1, SORT_LOOKBACK)])
def test_bucket_sort_size_limit(gpu, oracle, n, mode):
    # 2^19 slots is the largest one-view frame the bucket sort takes (buckets
    # of about 2065 keys); one splat more pads the view to 2^19 + 4096 slots
    # and the radix passes sort it. Both orders are the oracle's.
    rng = np.random.default_rng(9191)
    scene = synth.gaussian_scene(rng, n, 0, synth.mixture_means(rng, n))
    cam = gpu.look_at((0.0, 0.0, -4.0), (0.0, 0.0, 0.0))
    say(f"[bucket size limit] {n} splats")
    ctx = frame(gpu, scene, cam, 320, 240)
    assert ctx.stage_times().depth_sort == mode
    osc = oracle_scene(oracle, scene)
    keys_ref, idx_ref = oracle.preprocess(osc, cam_dict(cam))
    assert np.array_equal(ctx.sorted_indices(scene.count), oracle.sort_values_by_keys(keys_ref, idx_ref))
    ctx.close()
