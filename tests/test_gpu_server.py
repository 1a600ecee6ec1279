"""Served renders (gsplat_server_*) through the C ABI.

The bar: every request's image and visible count are exactly those of the
same view rendered alone (a one-view batch), whatever the batch that carried
it; requests arriving while batches run are folded into later batches;
memory admission refuses scenes and requests beyond the budget.
"""

import sys
import threading

import numpy as np
import pytest

import synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def alone(abi, dev, cam, w, h):
    b = abi.Batch()
    b.render([(dev, cam, w, h)])
    img = b.rgba8(0)
    vis = b.view_stats(0).visible_count
    b.close()
    return img, vis


def workload(abi, count, seed):
    specs = [(20000, 3), (5000, 0), (9000, 1), (40000, 3)]
    scenes = [synth.tiny_scene(n, 8100 + i, sh_degree=d) for i, (n, d) in enumerate(specs)]
    rng = np.random.default_rng(seed)
    sizes = [(320, 240), (640, 480), (256, 144), (97, 61)]
    reqs = []
    for _ in range(count):
        s = int(rng.integers(0, len(scenes)))
        w, h = sizes[int(rng.integers(0, len(sizes)))]
        reqs.append((s, abi.look_at(synth.sphere_point(rng, float(rng.uniform(2.0, 4.0)))), w, h))
    return scenes, reqs


@pytest.mark.parametrize("pipelines", [2, 1, 3])
def test_server_results_equal_alone(abi, pipelines):
    scenes, reqs = workload(abi, 48, 5)
    dev = [abi.Scene(s) for s in scenes]
    srv = abi.Server(max_batch_views=16, pipelines=pipelines)
    ids = [srv.add_scene(d) for d in dev]
    outs, tickets = [], []
    say(f"[server burst] {len(reqs)} requests, {pipelines} pipelines")
    for i, (s, cam, w, h) in enumerate(reqs):
        out = abi.PinnedImage(w, h) if i % 2 == 0 else np.zeros((h, w, 4), np.uint8)
        outs.append(out)
        tickets.append(srv.submit(ids[s], cam, w, h, out))
    results = [srv.wait(t) for t in tickets]
    st = srv.stats()
    assert st.completed == len(reqs) and st.rejected == 0
    assert st.batches < len(reqs), "a burst must be folded into shared batches"
    assert max(r.batch_views for r in results) > 1
    for i, ((s, cam, w, h), r, out) in enumerate(zip(reqs, results, outs)):
        img, vis = alone(abi, dev[s], cam, w, h)
        got = out.array if isinstance(out, abi.PinnedImage) else out
        assert r.status == 0 and r.visible_count == vis, f"request {i}"
        assert np.array_equal(got, img), f"request {i}: image differs from the view rendered alone"
        assert r.total_ms >= r.queue_ms >= 0.0
    srv.close()


def test_server_8x8_tiles_equal_default_alone(abi, monkeypatch):
    # Server pipelines rasterising in 8x8 tiles (raster_variant 4, read at
    # server creation) writing pinned and plain targets: every image equals
    # the view rendered alone with the default 16x16 raster.
    scenes, reqs = workload(abi, 24, 4)
    dev = [abi.Scene(s) for s in scenes]
    monkeypatch.setenv("GSPLAT_RASTER_VARIANT", "4")
    srv = abi.Server(max_batch_views=16, pipelines=2)
    monkeypatch.delenv("GSPLAT_RASTER_VARIANT")
    ids = [srv.add_scene(d) for d in dev]
    outs, tickets = [], []
    say(f"[server 8x8 tiles] {len(reqs)} requests")
    for i, (s, cam, w, h) in enumerate(reqs):
        out = abi.PinnedImage(w, h) if i % 2 == 0 else np.zeros((h, w, 4), np.uint8)
        outs.append(out)
        tickets.append(srv.submit(ids[s], cam, w, h, out))
    results = [srv.wait(t) for t in tickets]
    for i, ((s, cam, w, h), r, out) in enumerate(zip(reqs, results, outs)):
        img, vis = alone(abi, dev[s], cam, w, h)
        got = out.array if isinstance(out, abi.PinnedImage) else out
        assert r.status == 0 and r.visible_count == vis, f"request {i}"
        assert np.array_equal(got, img), f"request {i}: 8x8-tile image differs from the default raster alone"
    srv.close()


def test_server_concurrent_submitters(abi):
    scenes, reqs = workload(abi, 64, 6)
    dev = [abi.Scene(s) for s in scenes]
    srv = abi.Server(max_batch_views=64)
    ids = [srv.add_scene(d) for d in dev]
    results = [None] * len(reqs)
    outs = [abi.PinnedImage(w, h) for (_, _, w, h) in reqs]

    def client(k):
        for i in range(k, len(reqs), 8):
            s, cam, w, h = reqs[i]
            t = srv.submit(ids[s], cam, w, h, outs[i])
            results[i] = srv.wait(t)

    say("[server 8 client threads]")
    th = [threading.Thread(target=client, args=(k,)) for k in range(8)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for i, (s, cam, w, h) in enumerate(reqs[:16]):
        img, vis = alone(abi, dev[s], cam, w, h)
        assert results[i].visible_count == vis
        assert np.array_equal(outs[i].array, img), f"request {i}"
    assert srv.stats().completed == len(reqs)
    srv.close()


def test_server_memory_admission(abi):
    small = abi.Scene(synth.tiny_scene(1000, 11))
    big = abi.Scene(synth.tiny_scene(200000, 12))
    # a budget that holds the small scene and a 1080p working set per pipeline
    # (about 60 MB each, three pipelines) but not the big one (about 85 MB each)
    srv = abi.Server(memory_budget_bytes=220 << 20)
    sid = srv.add_scene(small)
    with pytest.raises(abi.GsplatError) as e:
        srv.add_scene(big)
    assert e.value.code == abi.UNSUPPORTED and "memory budget" in abi.last_error()
    # a request whose working set exceeds the pipeline share is refused, the next one runs
    huge = srv.submit(sid, abi.look_at((0, 0, -3)), 16384, 16384)
    ok = srv.submit(sid, abi.look_at((0, 0, -3)), 64, 64)
    r = srv.wait(huge, raise_on_error=False)
    assert r.status == abi.UNSUPPORTED
    assert srv.wait(ok).status == 0
    assert srv.stats().rejected == 1
    srv.close()


def test_server_errors(abi):
    dev = abi.Scene(synth.tiny_scene(100, 13))
    srv = abi.Server()
    sid = srv.add_scene(dev)
    cam = abi.look_at((0, 0, -3))
    for bad_call, code in [
        (lambda: srv.submit(sid + 100, cam, 64, 64), abi.INVALID_ARGUMENT),
        (lambda: srv.submit(sid, cam, 0, 64), abi.INVALID_ARGUMENT),
        (lambda: srv.submit(sid, cam, 64, 64, np.zeros(10, np.uint8)), abi.INVALID_ARGUMENT),
        (lambda: srv.wait(987654321), abi.INVALID_ARGUMENT),
        (lambda: srv.remove_scene(sid + 100), abi.INVALID_ARGUMENT),
    ]:
        with pytest.raises(abi.GsplatError) as e:
            bad_call()
        assert e.value.code == code
    badcam = abi.look_at((0, 0, -3))
    badcam.vertical_fov_radians = 0.0
    with pytest.raises(abi.GsplatError) as e:
        srv.submit(sid, badcam, 64, 64)
    assert e.value.code == abi.INVALID_ARGUMENT
    with pytest.raises(abi.GsplatError):
        abi.Server(max_batch_views=0)
    for bad in (0, 9):
        with pytest.raises(abi.GsplatError) as e:
            abi.Server(pipelines=bad)
        assert e.value.code == abi.INVALID_ARGUMENT
    # a removed scene no longer takes requests; queued ones still finish
    t = srv.submit(sid, cam, 64, 64)
    srv.remove_scene(sid)
    assert srv.wait(t).status == 0
    with pytest.raises(abi.GsplatError):
        srv.submit(sid, cam, 64, 64)
    srv.close()


def test_pinned_target_held_until_wait(abi):
    # gsplat_host_free refuses a block that a queued or running request targets,
    # and frees it once the request has ended; the binding keeps a target it was
    # handed alive until wait() even when the caller dropped it.
    import gc

    scene = synth.tiny_scene(60000, 21, sh_degree=3)
    dev = abi.Scene(scene)
    srv = abi.Server(max_batch_views=8)
    sid = srv.add_scene(dev)
    rng = np.random.default_rng(3)
    targets, tickets = [], []
    say("[server pinned hold] 120 requests at 1280x720")
    for _ in range(120):
        out = abi.PinnedImage(1280, 720)
        targets.append(out)
        tickets.append(srv.submit(sid, abi.look_at(synth.sphere_point(rng, 3.0)), 1280, 720, out))
    last = targets[-1]
    rc = abi.lib().gsplat_host_free(last.ptr)
    assert rc == abi.INVALID_ARGUMENT and "queued or running" in abi.last_error()
    with pytest.raises(abi.GsplatError):
        last.close()
    assert last.ptr  # still allocated
    dropped = srv.submit(sid, abi.look_at((0.2, 0.1, -3.0)), 320, 240, abi.PinnedImage(320, 240))
    gc.collect()
    for t in tickets:
        assert srv.wait(t).status == 0
    assert srv.wait(dropped).status == 0
    assert not srv._targets
    ref, _ = alone(abi, dev, abi.look_at(synth.sphere_point(np.random.default_rng(3), 3.0)), 1280, 720)
    assert np.array_equal(targets[0].array, ref)
    last.close()  # the request has ended
    assert last.ptr is None
    bogus = np.zeros(16, np.uint8)
    assert abi.lib().gsplat_host_free(bogus.ctypes.data) == abi.INVALID_ARGUMENT
    srv.close()


def test_removed_scene_counts_until_released(abi):
    # A removed scene's HBM stays counted against the budget while a pipeline's
    # last frame (or a queued request) still holds it, and is released after.
    a, b = abi.Scene(synth.tiny_scene(30000, 31)), abi.Scene(synth.tiny_scene(20000, 32))
    srv = abi.Server(pipelines=1)
    ia, ib = srv.add_scene(a), srv.add_scene(b)
    both = srv.stats().resident_scene_bytes
    assert both == a.device_bytes + b.device_bytes
    assert srv.wait(srv.submit(ia, abi.look_at((0, 0, -3)), 128, 96)).status == 0
    srv.remove_scene(ia)
    bytes_b = b.device_bytes
    a.close()  # the caller's handle may go too
    assert srv.stats().resident_scene_bytes == both and srv.stats().scenes == 1
    assert srv.wait(srv.submit(ib, abi.look_at((0, 0, -3)), 128, 96)).status == 0
    assert srv.stats().resident_scene_bytes == bytes_b
    srv.close()
