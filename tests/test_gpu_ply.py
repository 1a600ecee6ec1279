"""Real-scene ingest on the GPU: the 1080p workload's scene (300k splats, SH3)
written as a binary_little_endian PLY and loaded through
gsplat_context_load_scene_path renders bit-identically to the same scene
loaded from memory, and so does a point-count tier of it made by the
midpoint ladder (against the in-memory midpoint subset), which is also
checked against the oracle."""

import sys

import numpy as np
import pytest

import ply_tools
import synth
from helpers import cam_dict, oracle_scene

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def frame(abi, cam, w, h, scene=None, path=None):
    ctx = abi.Context(w, h)
    if path is not None:
        ctx.load_path(path)
    else:
        ctx.load(scene)
    ctx.set_camera(cam)
    ctx.render()
    n = ctx.stage_times().splat_count
    out = dict(order=ctx.sorted_indices(n), tiles=ctx.tile_counts(), img32=ctx.rgba32f(), img8=ctx.rgba8(),
               visible=ctx.stats().visible_count)
    ctx.close()
    return out


def same(a, b, label):
    assert a["visible"] == b["visible"], label
    assert np.array_equal(a["order"], b["order"]), f"{label}: permutation"
    assert np.array_equal(a["tiles"], b["tiles"]), f"{label}: tile counts"
    assert np.array_equal(a["img32"].view(np.uint32), b["img32"].view(np.uint32)), f"{label}: fp32 image"
    assert np.array_equal(a["img8"], b["img8"]), f"{label}: rgba8 image"


def test_binary_ply_scene_and_ladder_tier_render_as_loaded_from_memory(abi, oracle, tmp_path):
    scene, view = synth.plain_1080p()
    cam = abi.look_at(view.position, view.target)
    w, h = view.width, view.height
    src = ply_tools.write_scene_ply(tmp_path / "plain-1080p-300k-sh3.ply", scene)
    say(f"[ply] {src.stat().st_size} bytes written; rendering from path and from memory")
    same(frame(abi, cam, w, h, path=src), frame(abi, cam, w, h, scene=scene), "300k from PLY")
    m = ply_tools.generate_ladder(src, [100_000], tmp_path / "ladder")
    tier_path = m["tiers"][0]["local_path"]
    idx = ply_tools.midpoint_index(np.arange(100_000), 100_000, scene.count)
    sub = synth.Scene(scene.positions[idx], scene.opacity[idx], scene.scale_xyz[idx], scene.rotation_xyzw[idx],
                      scene.color_dc[idx], scene.sh_degree, scene.sh_rest[idx])
    say("[ply] 100k ladder tier from path and from memory")
    got = frame(abi, cam, w, h, path=tier_path)
    same(got, frame(abi, cam, w, h, scene=sub), "100k tier")
    say("[ply] 100k tier against the oracle")
    osc, cd = oracle_scene(oracle, sub), cam_dict(cam)
    keys, ids = oracle.preprocess(osc, cd)
    assert got["visible"] == len(ids)
    assert np.array_equal(got["order"], oracle.sort_values_by_keys(keys, ids))
    assert np.array_equal(got["tiles"], oracle.tile_counts(oracle.project(osc, cd, w, h), w, h))
    ref32, _ = oracle.render(osc, cd, w, h)
    assert np.abs(got["img32"] - ref32).max() <= 2e-3
