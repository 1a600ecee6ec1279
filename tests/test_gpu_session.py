"""Session order reuse of a context on the GPU (gsplat_context_set_sort_interval).

Along a slow orbit of the 1080p workload's scene, a context with sort interval
2 refreshes its depth order on every second camera change. Refreshed frames
must equal the plain path's frame bit for bit; the frames between present
the current camera's projection in the previous refresh's order, which must
be exactly that refresh's order, and their RGBA8 images must pass the
reference's stale-order gate against the exact frame (rgba_image_parity_
metrics, render-wgpu/src/lib.rs:3491-3545: mean |delta RGB| <= 1/255 and at most
0.1% of pixels off by more than 3/255). A camera jump outside the pose
envelope refreshes at once (flag bit 7)."""

import math
import sys

import numpy as np
import pytest

import synth
from helpers import rgba_parity_metrics

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def orbit(abi, k, step=0.003, radius=4.0, base=0.7):
    a = base + step * k
    return abi.look_at((radius * math.cos(a), 0.8, radius * math.sin(a)))


def grab(ctx, n):
    return ctx.rgba8(), ctx.sorted_indices(n), ctx.rgba32f()


def test_sort_interval_two_reuses_orders_within_the_gate(abi):
    scene, view = synth.plain_1080p()
    w, h, n = view.width, view.height, scene.count
    exact, sess = abi.Context(w, h), abi.Context(w, h)
    for c in (exact, sess):
        c.load(scene)
    sess.set_sort_interval(2)
    last_refresh_order = None
    reused = refreshed = 0
    worst = (0.0, 0.0)
    for k in range(12):
        cam = orbit(abi, k)
        exact.set_camera(cam)
        exact.render()
        e8, eo, e32 = grab(exact, n)
        sess.set_camera(cam)
        sess.render()
        s8, so, s32 = grab(sess, n)
        st = sess.sort_stats()
        assert st.camera_revision == k + 1
        if st.flags & 1:
            refreshed += 1
            assert st.presented_order_revision_lag == 0 and st.applied_order_revision == k + 1
            assert np.array_equal(so, eo), f"frame {k}: refreshed order differs"
            assert np.array_equal(s32.view(np.uint32), e32.view(np.uint32)), f"frame {k}: refreshed fp32 image"
            assert np.array_equal(s8, e8), f"frame {k}: refreshed image"
            last_refresh_order = eo
        else:
            reused += 1
            assert st.presented_order_revision_lag == 1 and not st.flags & (1 << 7)
            assert np.array_equal(so, last_refresh_order), f"frame {k}: not the kept order"
            mean, frac, mx = rgba_parity_metrics(s8, e8)
            worst = (max(worst[0], mean), max(worst[1], frac))
            say(f"[session] frame {k} reused: mean {mean * 255:.4f}/255, over 3/255 {frac:.5f}, max {mx * 255:.1f}/255")
            assert mean <= 1.0 / 255.0 and frac <= 0.001, f"frame {k}: stale order outside the gate"
    assert (refreshed, reused) == (6, 6)
    # one more step refreshes (the second change); a jump beyond the pose
    # envelope right after it refreshes at once, on the first change
    for c in (exact, sess):
        c.set_camera(orbit(abi, 12))
        c.render()
    assert sess.sort_stats().flags & 1
    far = abi.look_at((0.0, 4.0, 0.3))
    for c in (exact, sess):
        c.set_camera(far)
        c.render()
    st = sess.sort_stats()
    assert st.flags & 1 and st.flags & (1 << 7)
    assert np.array_equal(sess.rgba8(), exact.rgba8())
    with pytest.raises(abi.GsplatError) as e:
        sess.set_sort_interval(0)
    assert e.value.code == abi.INVALID_ARGUMENT
    say(f"[session] worst reused frame: mean {worst[0] * 255:.4f}/255, over 3/255 {worst[1]:.5f}")
    exact.close()
    sess.close()


def test_unchanged_camera_reuses_its_own_order_bit_for_bit(abi):
    scene = synth.tiny_scene(20000, 91, sh_degree=3)
    ctx = abi.Context(320, 240)
    ctx.load(scene)
    ctx.set_sort_interval(1)
    ctx.set_camera(abi.look_at((0.3, 0.2, -2.8)))
    ctx.render()
    first = grab(ctx, scene.count)
    ctx.render()  # the same camera: the kept order is this camera's own
    assert not ctx.sort_stats().flags & 1
    again = grab(ctx, scene.count)
    for a, b in zip(first, again):
        assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))
    ctx.force_sort_refresh()
    ctx.render()
    assert ctx.sort_stats().flags & 1
    ctx.close()
