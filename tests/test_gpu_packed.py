"""The packed geometry path (GSPLAT_GEOMETRY_PATH_PACKED_ATLAS) on the GPU.

Bars:
- the reference's packed-vs-direct gates (render-wgpu/src/lib.rs:3219-3229,
  3733-3771, 3964-3976): equal visible and drawn counts, and RGBA8 frames
  within mean |dRGB| <= 1/255 with at most 0.1% of pixels off by more than
  3/255, on its minimal two-splat scene (64 and 128 px), minimal_ascii.ply
  and its 64-splat degree-3 scene (128 px); here also on the 16k SH0 640x480
  and 300k SH3 1080p workloads;
- decode parity: the packed projection of every splat equals the oracle's
  projection of the scene the CPU decoders say the records hold
  (tests/helpers.py decode_packed), to float tolerance (the codec's
  divisions are reciprocal multiplications on the device);
- batches and served requests mixing packed and direct scenes give, view for
  view, exactly the frame a context renders in the same path.
"""

import pathlib
import sys

import numpy as np
import pytest

import synth
from helpers import cam_dict, packed_decoded_scene, rgba_parity_metrics

pytestmark = pytest.mark.gpu
GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
DIRECT, PACKED = 0, 1


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def build_scene():
    """build_scene of render-wgpu/src/lib.rs:3072-3082."""
    return synth.Scene(np.float32([[0, 0, 0.5], [0, 0, 2.0]]), np.float32([0.9, 0.8]),
                       np.float32([[0, 0, 0], [0.2, 0.2, 0.2]]), np.float32([[0, 0, 0, 1], [0, 0, 0, 1]]),
                       np.float32([[0.1, 0.2, 0.3], [0.3, 0.2, 0.1]]), 0, None)


def synthetic_degree3_scene():
    """synthetic_degree3_scene of render-wgpu/src/lib.rs:3917-3939."""
    n = 64
    i = np.arange(n)
    t = (i / n).astype(np.float32)
    pos = np.stack([(t - np.float32(0.5)) * np.float32(2.0), np.zeros(n, np.float32),
                    np.float32(1.5) + (i // 16).astype(np.float32) * np.float32(0.1)], 1)
    dc = np.stack([np.float32(0.1) + (i % 5).astype(np.float32) * np.float32(0.05), np.full(n, -0.05, np.float32),
                   np.full(n, 0.2, np.float32)], 1)
    rest = (((np.arange(n * 45) % 11).astype(np.float32) - np.float32(5.0)) * np.float32(0.03)).reshape(n, 45)
    return synth.Scene(pos.astype(np.float32), np.full(n, 2.0, np.float32), np.full((n, 3), -3.0, np.float32),
                       np.tile(np.float32([0, 0, 0, 1]), (n, 1)), dc.astype(np.float32), 3, rest)


def render(abi, scene, cam, w, h, path, ply=None):
    ctx = abi.Context(w, h)
    ctx.set_geometry_path(path)
    if ply is not None:
        ctx.load_path(ply)
    else:
        ctx.load(scene)
    ctx.set_camera(cam)
    ctx.render()
    return ctx


def gate(abi, scene, cam, w, h, label, ply=None):
    say(f"[{label}] packed vs direct {w}x{h}")
    a = render(abi, scene, cam, w, h, DIRECT, ply)
    b = render(abi, scene, cam, w, h, PACKED, ply)
    assert b.geometry_path() == PACKED and a.geometry_path() == DIRECT
    sa, sb = a.stats(), b.stats()
    mean, frac, worst = rgba_parity_metrics(a.rgba8(), b.rgba8())
    say(f"[{label}] mean_abs_rgb={mean:.6f} frac_over_3_255={frac:.6f} max_abs_rgb={worst:.6f} "
        f"visible {sa.visible_count}/{sb.visible_count}")
    assert mean <= 1.0 / 255.0, f"{label}: mean abs RGB {mean:.6f} exceeds 1/255"
    assert frac <= 0.001, f"{label}: {frac:.6f} of pixels off by more than 3/255"
    out = (sa, sb)
    a.close()
    b.close()
    return out


def test_packed_counts_on_minimal_scene(abi):
    cam = abi.lib().gsplat_camera_default()
    sa, sb = gate(abi, build_scene(), cam, 64, 64, "minimal 64")
    assert sa.visible_count == sb.visible_count == 2
    assert sa.drawn_count == sb.drawn_count == 2


@pytest.mark.parametrize("case", ["build_scene", "minimal_ascii", "synthetic_degree3"])
def test_packed_vs_direct_image_gate(abi, case):
    cam = abi.lib().gsplat_camera_default()
    if case == "minimal_ascii":
        sa, sb = gate(abi, None, cam, 128, 128, case, ply=GOLDEN / "minimal_ascii.ply")
    else:
        sa, sb = gate(abi, build_scene() if case == "build_scene" else synthetic_degree3_scene(), cam, 128, 128, case)
    if case != "synthetic_degree3":  # the reference checks counts on these two
        assert sa.visible_count == sb.visible_count and sa.drawn_count == sb.drawn_count


@pytest.mark.parametrize("workload", ["plain-640x480-16k-sh0", "plain-1080p-300k-sh3"])
def test_packed_vs_direct_gate_at_workload_size(abi, workload):
    scene, view = synth.plain_small() if workload.startswith("plain-640") else synth.plain_1080p()
    cam = abi.look_at(view.position, view.target)
    sa, sb = gate(abi, scene, cam, view.width, view.height, workload)
    assert abs(int(sa.visible_count) - int(sb.visible_count)) <= max(2, sa.visible_count // 10000)
    d = abi.Scene(scene)
    p = abi.Scene(scene, geometry_path=PACKED)
    per = 17 if scene.sh_degree else 5
    assert p.geometry_path == PACKED and d.geometry_path == DIRECT
    assert p.device_bytes == 4 * (per * scene.count + abi.PACKED_CONSTS)
    assert d.device_bytes == 4 * (14 + 3 * ((scene.sh_degree + 1) ** 2 - 1)) * scene.count
    d.close()
    p.close()


@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_packed_decode_matches_oracle_of_decoded_scene(abi, oracle, degree):
    scene = synth.tiny_scene(3000, 9100 + degree, sh_degree=degree)
    cam = abi.look_at((0.4, -0.3, -2.5))
    w, h = 320, 240
    say(f"[decode] degree {degree} N={scene.count}")
    ctx = render(abi, scene, cam, w, h, PACKED)
    got = ctx.projection(scene.count)
    keys = ctx.depth_keys(scene.count)
    ctx.close()
    words, consts = abi.pack_scene(scene)
    osc = packed_decoded_scene(oracle, words, consts, degree)
    ref = oracle.project(osc, cam_dict(cam), w, h)
    keys_ref, idx_ref = oracle.preprocess(osc, cam_dict(cam))
    exp_keys = np.zeros(scene.count, np.uint32)
    exp_keys[idx_ref] = keys_ref
    # depth: the same splats in front, depths within a few ulps
    assert np.array_equal(keys != 0, exp_keys != 0)
    assert np.allclose(keys.view(np.float32), exp_keys.view(np.float32), rtol=2e-6, atol=0)
    drawn = (got[:, 0] == 1) & (ref[:, 0] == 1)
    assert ((got[:, 0] == 1) != (ref[:, 0] == 1)).sum() <= 2
    assert drawn.sum() > scene.count // 2
    g, r = got[drawn], ref[drawn]
    assert np.abs(g[:, 1:3] - r[:, 1:3]).max() < 1e-5, "centres"

    def form(p):  # the ellipse's quadratic form [u v][u v]^T, stable where the axes are not
        ux, uy, vx, vy = p[:, 3], p[:, 4], p[:, 5], p[:, 6]
        return np.stack([ux * ux + vx * vx, ux * uy + vx * vy, uy * uy + vy * vy], 1)

    fg, fr = form(g), form(r)
    scale = np.abs(fr).max(1, keepdims=True)
    assert (np.abs(fg - fr) / scale).max() < 1e-4, "ellipse forms"
    assert np.abs(g[:, 7:11] - r[:, 7:11]).max() < 1e-5, "premultiplied colour and alpha"


def test_mixed_batch_and_server_equal_context(abi):
    specs = [(5000, 0, PACKED), (12000, 3, PACKED), (4097, 1, PACKED), (9000, 2, PACKED), (3000, 4, DIRECT),
             (7000, 3, DIRECT)]
    scenes = [synth.tiny_scene(n, 9200 + i, sh_degree=d) for i, (n, d, _) in enumerate(specs)]
    dev = [abi.Scene(s, geometry_path=p) for s, (_, _, p) in zip(scenes, specs)]
    rng = np.random.default_rng(5)
    sizes = [(320, 240), (257, 16), (640, 480), (97, 61)]
    views = []
    for v in range(14):  # two or three views per scene: projection groups of one scene
        s = v % len(scenes)
        w, h = sizes[v % len(sizes)]
        views.append((s, abi.look_at(synth.sphere_point(rng, float(rng.uniform(2.0, 4.0)))), w, h))
    say(f"[mixed packed batch] {len(views)} views")
    b = abi.Batch()
    b.render([(dev[s], cam, w, h) for (s, cam, w, h) in views])
    srv = abi.Server(max_batch_views=16)
    ids = [srv.add_scene(d) for d in dev]
    outs = [np.zeros((h, w, 4), np.uint8) for (_, _, w, h) in views]
    tickets = [srv.submit(ids[s], cam, w, h, out) for (s, cam, w, h), out in zip(views, outs)]
    results = [srv.wait(t) for t in tickets]
    for v, (s, cam, w, h) in enumerate(views):
        ctx = render(abi, scenes[s], cam, w, h, specs[s][2])
        img8 = ctx.rgba8()
        img32 = ctx.rgba32f()
        vis = ctx.stats().visible_count
        order = ctx.sorted_indices(scenes[s].count)
        ctx.close()
        label = f"view {v} (scene {s}, path {specs[s][2]})"
        assert b.view_stats(v).visible_count == vis, label
        assert np.array_equal(b.sorted_indices(v, scenes[s].count), order), f"{label}: permutation"
        assert np.array_equal(b.rgba32f(v).view(np.uint32), img32.view(np.uint32)), f"{label}: fp32 image"
        assert np.array_equal(b.rgba8(v), img8), f"{label}: rgba8 image"
        assert results[v].status == 0 and results[v].visible_count == vis, f"{label}: served"
        assert np.array_equal(outs[v], img8), f"{label}: served image"
    srv.close()
    b.close()


def test_context_geometry_path_switch(abi):
    scene = synth.tiny_scene(6000, 9300, sh_degree=3)
    cam = abi.look_at((0.0, 0.5, -3.0))
    w, h = 256, 192

    def frame(ctx):
        ctx.set_camera(cam)
        ctx.render()
        return ctx.rgba8()

    images = []
    for path in (DIRECT, PACKED):
        c = render(abi, scene, cam, w, h, path)
        images.append(c.rgba8())
        c.close()
    direct, packed = images
    ctx = abi.Context(w, h)
    ctx.load(scene)
    assert ctx.geometry_path() == DIRECT
    assert np.array_equal(frame(ctx), direct)
    ctx.set_geometry_path(PACKED)
    assert ctx.geometry_path() == PACKED and np.array_equal(frame(ctx), packed)
    ctx.set_geometry_path(DIRECT)
    assert np.array_equal(frame(ctx), direct)
    for bad, code in ((2, abi.UNSUPPORTED), (3, abi.INVALID_ARGUMENT), (0xFFFFFFFF, abi.INVALID_ARGUMENT)):
        with pytest.raises(abi.GsplatError) as e:
            ctx.set_geometry_path(bad)
        assert e.value.code == code
        assert ctx.geometry_path() == DIRECT
    # SH degree 4 has no packed form: refused, path and scene kept.
    ctx4 = abi.Context(w, h)
    s4 = synth.tiny_scene(2000, 9301, sh_degree=4)
    ctx4.load(s4)
    before = frame(ctx4)
    with pytest.raises(abi.GsplatError) as e:
        ctx4.set_geometry_path(PACKED)
    assert e.value.code == abi.UNSUPPORTED and ctx4.geometry_path() == DIRECT
    assert np.array_equal(frame(ctx4), before)
    with pytest.raises(abi.GsplatError) as e:
        abi.Scene(s4, geometry_path=PACKED)
    assert e.value.code == abi.UNSUPPORTED
    ctx4.close()
    ctx.close()
