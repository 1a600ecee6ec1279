"""Pins the CPU oracle against the reference's own known answers and against
closed forms / brute force (the oracle is the checker of the HIP path)."""

import math

import numpy as np
import pytest

from helpers import DEFAULT_FOV, default_cam, lcg_keys, lcg_keys_fast, minimal_ascii_scene


# ---------------------------------------------------------------- sort KATs

def test_sort_values_by_immutable_keys(oracle):
    # sort/src/lib.rs:678-688
    assert list(oracle.sort_values_by_keys([20, 5, 12, 12], [0, 1, 2, 3])) == [0, 2, 3, 1]


def test_sort_descending_three(oracle):
    # sort/src/lib.rs:602-612 (values start ascending, so value sort == pair sort)
    assert list(oracle.sort_values_by_keys([20, 5, 12], [0, 1, 2])) == [0, 2, 1]


@pytest.mark.parametrize("length,seed", [(0, 42), (1, 43), (2, 44), (255, 297), (256, 298), (257, 299),
                                         (1024, 1066), (4099, 1), (65537, 42 + 65537)])
def test_sort_matches_reference_order(oracle, length, seed):
    # sort/src/lib.rs:614-647: expected = sort_by(desc key, asc index)
    keys = lcg_keys_fast(length, seed)
    got = oracle.sort_values_by_keys(keys, np.arange(length, dtype=np.uint32))
    expected = np.lexsort((np.arange(length), -keys.astype(np.int64))) if length else np.zeros(0, np.int64)
    assert np.array_equal(got, expected.astype(np.uint32))


def test_lcg_fast_matches_scalar():
    assert np.array_equal(lcg_keys_fast(5000, 99), lcg_keys(5000, 99))


def test_sort_is_stable_on_heavy_ties(oracle):
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 4, 10000).astype(np.uint32)
    vals = rng.permutation(10000).astype(np.uint32)
    got = oracle.sort_values_by_keys(keys, vals)
    order = np.argsort(-keys.astype(np.int64), kind="stable")
    assert np.array_equal(got, vals[order])


# ------------------------------------------------------ camera / preprocess

def test_quat_inverse_normalizes_scaled_input(oracle):
    # render-wgpu/src/lib.rs:4113-4116: quat_inverse([0,0,0,2]) == identity
    assert np.array_equal(oracle.view_rows([0, 0, 0, 2]), np.eye(3, dtype=np.float32))


def test_view_rows_are_orthonormal_and_invert_rotation(oracle):
    rng = np.random.default_rng(1)
    for _ in range(20):
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        r = oracle.view_rows(q).astype(np.float64)
        assert np.allclose(r @ r.T, np.eye(3), atol=1e-5)
        x, y, z, w = q  # rotation matrix of q; the view rows are its transpose
        m = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        assert np.allclose(r, m.T, atol=1e-5)


def test_camera_validation(oracle):
    c = default_cam()
    assert oracle.camera_valid(c)
    for key, val in [("near", 10.0), ("fov", 0.0), ("fov", math.pi), ("near", 0.0)]:
        bad = dict(c)
        bad[key] = val
        if key == "near" and val == 10.0:
            bad["far"] = 1.0
        assert not oracle.camera_valid(bad)
    assert not oracle.camera_valid(dict(c, rotation_xyzw=[0, 0, 0, 2]))
    assert oracle.camera_valid(dict(c, rotation_xyzw=[0, 0, 0, 1.0004]))


def test_back_to_front_order_with_culling(oracle):
    # render-wgpu/src/lib.rs:4062-4089: 5 splats, 3 visible, order [1, 2, 0]
    pos = [[0, 0, 0.5], [0, 0, 2.0], [0, 0, 1.0], [0, 0, -1.0], [0, 0, 2000.0]]
    s = oracle.Scene(pos, [1.0] * 5, [[0, 0, 0]] * 5, [[0, 0, 0, 1]] * 5, [[0.2, 0.3, 0.4]] * 5)
    keys, idx = oracle.preprocess(s, default_cam())
    assert list(idx) == [0, 1, 2]
    order = oracle.sort_values_by_keys(keys, idx)
    assert list(order) == [1, 2, 0]


def test_depth_keys_closed_form(oracle):
    rng = np.random.default_rng(3)
    n = 2000
    pos = rng.normal(0, 2, (n, 3)).astype(np.float32)
    s = oracle.Scene(pos, np.zeros(n), np.zeros((n, 3)), np.tile([0, 0, 0, 1], (n, 1)), np.zeros((n, 3)))
    cam = dict(default_cam(), position=[0.1, -0.2, -3.0])
    keys, idx = oracle.preprocess(s, cam)
    # identity rotation: depth = 1*(z-cz) + 0*(x-cx) + 0*(y-cy), exact in f32
    d = (pos[:, 2] - np.float32(-3.0)).astype(np.float32)
    vis = (d >= np.float32(0.01)) & (d <= np.float32(1000.0))
    assert np.array_equal(idx, np.nonzero(vis)[0].astype(np.uint32))
    assert np.array_equal(keys, d[vis].view(np.uint32))


# ------------------------------------------------------ precompute / project

def test_precompute_closed_forms(oracle):
    qz = np.float32(math.sqrt(0.5))
    s = oracle.Scene([[0, 0, 2], [0, 0, 2]], [0.0, 2.0], [[0.8, -0.4, -0.4], [0.1, 0.2, 0.3]],
                     [[0, 0, qz, qz], [0, 0, 0, 1]], [[0, 0, 0]] * 2)
    alpha, cov = oracle.precompute(s)
    assert alpha[0] == np.float32(0.5)
    assert abs(alpha[1] - 1 / (1 + math.exp(-2))) < 1e-7
    # identity rotation: diag(e^{2s})
    assert np.allclose(cov[1], [math.exp(0.2), 0, 0, math.exp(0.4), 0, math.exp(0.6)], rtol=1e-6, atol=1e-9)
    # 90 degrees about z swaps the x and y variances
    assert np.allclose(cov[0], [math.exp(-0.8), 0, 0, math.exp(1.6), 0, math.exp(-0.8)], rtol=1e-5, atol=1e-6)


def test_projection_axes_anisotropic_orthogonal(oracle):
    # render-wgpu/src/lib.rs:4463-4496
    qz = np.float32(math.sqrt(0.5))
    s = oracle.Scene([[0, 0, 2]], [1.0], [[0.8, -0.4, -0.4]], [[0, 0, qz, qz]], [[0.2, 0.3, 0.4]])
    p = oracle.project(s, default_cam(), 1280, 720)[0]
    lu, lv = np.hypot(p[3], p[4]), np.hypot(p[5], p[6])
    assert p[0] == 1 and lu > lv and abs(p[3] * p[5] + p[4] * p[6]) < 1e-4


def test_projection_keeps_partial_offscreen_splat(oracle):
    # render-wgpu/src/lib.rs:4498-4528
    s = oracle.Scene([[2.3, 0, 1]], [1.0], [[2, 2, 2]], [[0, 0, 0, 1]], [[0.2, 0.3, 0.4]])
    p = oracle.project(s, default_cam(), 1280, 720)[0]
    assert p[0] == 1 and p[1] > 2.0 and p[1] - (abs(p[3]) + abs(p[5])) <= 1.0


def test_projection_isotropic_closed_form(oracle):
    # Isotropic splat on the axis: center 0, axes 3*sqrt(f^2 s^2/z^2 + blur) per axis (clamped).
    s_log = -3.0
    z = 2.0
    s = oracle.Scene([[0, 0, z]], [0.0], [[s_log] * 3], [[0, 0, 0, 1]], [[0.1, 0.2, 0.3]])
    w, h = 800, 800
    p = oracle.project(s, default_cam(), w, h)[0]
    f = 1 / math.tan(float(DEFAULT_FOV) / 2)
    var = (f * math.exp(s_log) / z) ** 2 + (0.3 * 2 / w) ** 2
    r = 3 * math.sqrt(var)
    assert p[0] == 1 and p[1] == 0 and p[2] == 0
    assert abs(math.hypot(p[3], p[4]) - r) < 1e-6 and abs(math.hypot(p[5], p[6]) - r) < 1e-6
    # colour: clamp(C0*dc + 0.5) * alpha with alpha = sigmoid(0) = 0.5
    c0 = 0.28209479177387814
    assert np.allclose(p[7:11], [(c0 * 0.1 + 0.5) * 0.5, (c0 * 0.2 + 0.5) * 0.5, (c0 * 0.3 + 0.5) * 0.5, 0.5],
                       atol=1e-7)


def _real_sh_basis(deg, d):
    """Independent real SH basis (scipy associated Legendre), 3DGS sign convention, k = 1..(deg+1)^2-1."""
    from scipy.special import lpmv
    x, y, z = d
    theta = math.acos(max(-1.0, min(1.0, z)))
    phi = math.atan2(y, x)
    out = []
    for l in range(1, deg + 1):
        for m in range(-l, l + 1):
            am = abs(m)
            norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
            p = lpmv(am, l, math.cos(theta)) * (-1) ** am  # remove Condon-Shortley phase
            if m == 0:
                v = norm * p
            elif m > 0:
                v = math.sqrt(2) * norm * p * math.cos(m * phi)
            else:
                v = math.sqrt(2) * norm * p * math.sin(am * phi)
            out.append(v)
    return np.array(out)


@pytest.mark.parametrize("deg", [1, 2, 3, 4])
def test_sh_colour_matches_independent_basis(oracle, deg):
    # eval_sh_rgb (wgsl:233-326) uses 3DGS's real SH with signs (-1)^... on odd-m terms;
    # compare |Y| and the colour for unit coefficient vectors.
    rng = np.random.default_rng(deg)
    k = (deg + 1) ** 2 - 1
    for trial in range(6):
        pos = rng.normal(0, 1, 3).astype(np.float32)
        pos[2] = abs(pos[2]) + 1.0
        d = pos / np.linalg.norm(pos)
        basis = _real_sh_basis(deg, d)
        for c in range(k):
            rest = np.zeros((1, 3 * k), np.float32)
            rest[0, c] = 0.1
            s = oracle.Scene([pos], [30.0], [[-3, -3, -3]], [[0, 0, 0, 1]], [[0, 0, 0]], deg, rest)
            p = oracle.project(s, default_cam(), 64, 64)[0]
            if p[0] == 0:
                continue
            val = (p[7] - 0.5)  # alpha == 1 (sigmoid(30) rounds to 1), red channel
            assert abs(abs(val) - abs(0.1 * basis[c])) < 2e-6, (deg, c)


def test_coverage_cut_is_the_last_passing_r2(oracle):
    import ctypes
    libm = ctypes.CDLL("libm.so.6")  # the fragment test is stated with the C library's expf
    libm.expf.restype = ctypes.c_float
    libm.expf.argtypes = [ctypes.c_float]
    for alpha in [0.001, 1 / 256, 0.0040, 0.01, 0.1, 0.3, 0.352, 0.36, 0.9, 1.0]:
        cut = oracle.coverage_cut(alpha)
        a = np.float32(alpha)
        passes = lambda r2: r2 <= 1.0 and a * np.float32(libm.expf(np.float32(-4.5) * np.float32(r2))) > np.float32(1 / 256)  # noqa
        if cut < 0:
            assert not passes(np.float32(0.0))
            continue
        c = np.float32(cut)
        assert passes(c)
        if c < 1.0:
            assert not passes(np.nextafter(c, np.float32(2.0)))


# ------------------------------------------------------------- raster KATs

def test_conformance_means(oracle):
    # crates/gsplat-render-wgpu/tests/conformance_sorted_alpha.rs:43-76
    img, vis = oracle.render(minimal_ascii_scene(oracle), default_cam(), 64, 64)
    assert vis == 2
    u8 = oracle.to_rgba8(img)
    assert (u8[..., 3] > 0).any()
    means = u8.reshape(-1, 4).astype(np.float64).mean(0) / 255.0
    assert np.all(np.abs(means - np.array([0.4194, 0.3234, 0.3073, 0.5712])) <= 0.035), means


def test_single_splat_analytic_image(oracle):
    # One isotropic splat at the image centre: pixel = rgb*alpha*g over transparent black.
    z, s_log, w = 2.0, -2.5, 96
    s = oracle.Scene([[0, 0, z]], [1.0], [[s_log] * 3], [[0, 0, 0, 1]], [[0.3, -0.2, 0.1]])
    img, vis = oracle.render(s, default_cam(), w, w)
    p = oracle.project(s, default_cam(), w, w)[0]
    R = math.hypot(p[3], p[4])
    a = 1 / (1 + math.exp(-1.0))
    c0 = 0.28209479177387814
    rgb = [min(max(c0 * v + 0.5, 0), 1) * a for v in (0.3, -0.2, 0.1)]
    yy, xx = np.mgrid[0:w, 0:w]
    pnx = 2 * (xx + 0.5) / w - 1
    pny = 1 - 2 * (yy + 0.5) / w
    r2 = (pnx ** 2 + pny ** 2) / R ** 2
    g = np.exp(-4.5 * r2)
    cover = (r2 <= 1) & (a * g > 1 / 256)
    exp = np.zeros((w, w, 4))
    for ch in range(3):
        exp[..., ch] = np.where(cover, rgb[ch] * g, 0)
    exp[..., 3] = np.where(cover, a * g, 0)
    border = np.abs(np.sqrt(r2) - math.sqrt(min(1.0, math.log(256 * a) / 4.5))) < 1e-3
    assert np.abs(img - exp)[~border].max() < 1e-5


@pytest.mark.parametrize("cx,cy", [(0.0, 0.0), (-0.148, 0.31), (0.5, -0.62), (0.853, 0.853)])
def test_per_pixel_quadratic_form_matches_float64(oracle, cx, cy):
    # The oracle's per-pixel r^2 (evaluated relative to 128-pixel block
    # origins, centre terms folded first) against |A^-1 (p - c)|^2 in
    # float64 from the same projected axes: small anisotropic splats placed
    # on and across block boundaries of a 300x300 frame (blocks at 0, 128, 256).
    w = 300
    z = 2.0
    f = 1.0 / math.tan(float(DEFAULT_FOV) / 2)
    s = oracle.Scene([[cx * z / f, cy * z / f, z]], [2.0], [[-4.2, -5.0, -4.6]], [[0.2, -0.4, 0.3, 0.85]],
                     [[0.3, -0.2, 0.1]])
    img, vis = oracle.render(s, default_cam(), w, w)
    p = oracle.project(s, default_cam(), w, w)[0].astype(np.float64)
    assert vis == 1 and p[0] == 1
    A = np.array([[p[3], p[5]], [p[4], p[6]]])  # columns axis_u, axis_v
    Ai = np.linalg.inv(A)
    yy, xx = np.mgrid[0:w, 0:w]
    d = np.stack([2 * (xx + 0.5) / w - 1 - p[1], 1 - 2 * (yy + 0.5) / w - p[2]], -1)
    l = d @ Ai.T
    r2 = (l ** 2).sum(-1)
    a = 1 / (1 + math.exp(-2.0))
    g = np.exp(-4.5 * r2)
    cover = (r2 <= 1) & (a * g > 1 / 256)
    border = np.abs(np.sqrt(r2) - math.sqrt(min(1.0, math.log(256 * a) / 4.5))) < 2e-3
    assert cover.sum() > 20
    got_a = img[..., 3]
    exp_a = np.where(cover, a * g, 0)
    assert np.abs(got_a - exp_a)[~border].max() < 2e-5


def test_render_window_equals_full_crop(oracle):
    # A window render computes each pixel exactly as the full render does.
    rng = np.random.default_rng(5)
    n = 300
    pos = np.c_[rng.normal(0, 0.4, (n, 2)), rng.uniform(1.5, 3.5, n)]
    s = oracle.Scene(pos, rng.normal(0, 1.5, n), rng.normal(-2.8, 0.5, (n, 3)),
                     rng.standard_normal((n, 4)) / 2, rng.normal(0, 0.3, (n, 3)))
    w, h = 97, 61
    full, vis = oracle.render(s, default_cam(), w, h)
    assert (full[..., 3] > 0).mean() > 0.2
    for x0, y0, x1, y1 in [(0, 0, w, h), (13, 7, 50, 33), (90, 0, 97, 1), (0, 60, 97, 61), (40, 20, 41, 21)]:
        win, v2 = oracle.render_window(s, default_cam(), w, h, x0, y0, x1, y1)
        assert v2 == vis
        assert np.array_equal(win, full[y0:y1, x0:x1]), (x0, y0, x1, y1)
    for bad in [(5, 5, 5, 9), (0, 0, w + 1, h), (3, 9, 8, 7)]:
        with pytest.raises(ValueError):
            oracle.render_window(s, default_cam(), w, h, *bad)
    # Load-time values computed once give the same frame, on any thread count.
    prepared = oracle.precompute(s)
    for threads in (1, 3, 8):
        img, v3 = oracle.render_prepared(s, prepared, default_cam(), w, h, threads=threads)
        assert v3 == vis and np.array_equal(img.view(np.uint32), full.view(np.uint32)), threads


def test_front_to_back_equals_back_to_front(oracle):
    # The oracle blends back to front; restate front to back in float64 and compare.
    rng = np.random.default_rng(11)
    n = 40
    pos = np.c_[rng.normal(0, 0.2, (n, 2)), rng.uniform(1.5, 3.0, n)]
    s = oracle.Scene(pos, rng.normal(0, 1.5, n), rng.normal(-2.5, 0.3, (n, 3)),
                     rng.standard_normal((n, 4)) / 2, rng.normal(0, 0.3, (n, 3)))
    w = 48
    img, vis = oracle.render(s, default_cam(), w, w)
    proj = oracle.project(s, default_cam(), w, w)
    keys, idx = oracle.preprocess(s, default_cam())
    order = oracle.sort_values_by_keys(keys, idx)[::-1]  # front to back
    yy, xx = np.mgrid[0:w, 0:w]
    pnx = (2 * (xx + 0.5) / w - 1).astype(np.float64)
    pny = (1 - 2 * (yy + 0.5) / w).astype(np.float64)
    C = np.zeros((w, w, 3))
    T = np.ones((w, w))
    for i in order:
        p = proj[i]
        if p[0] == 0:
            continue
        A = np.array([[p[3], p[5]], [p[4], p[6]]], np.float64)
        inv = np.linalg.inv(A)
        dx, dy = pnx - p[1], pny - p[2]
        lx = inv[0, 0] * dx + inv[0, 1] * dy
        ly = inv[1, 0] * dx + inv[1, 1] * dy
        r2 = lx * lx + ly * ly
        g = np.exp(-4.5 * r2)
        a = p[10] * g
        m = (r2 <= 1) & (a > 1 / 256)
        for ch in range(3):
            C[..., ch] += np.where(m, T * p[7 + ch] * g, 0)
        T = np.where(m, T * (1 - a), T)
    assert np.abs(img[..., :3] - C).max() < 2e-4
    assert np.abs(img[..., 3] - (1 - T)).max() < 2e-4


def test_tile_counts_cover_every_covered_pixel(oracle):
    rng = np.random.default_rng(7)
    n = 300
    pos = np.c_[rng.normal(0, 0.6, (n, 2)), rng.uniform(1.0, 4.0, n)]
    s = oracle.Scene(pos, rng.normal(0, 1.5, n), rng.normal(-3.0, 0.6, (n, 3)),
                     rng.standard_normal((n, 4)), rng.normal(0, 0.3, (n, 3)))
    w, h = 200, 120
    proj = oracle.project(s, default_cam(), w, h)
    yy, xx = np.mgrid[0:h, 0:w]
    pnx = (np.float32(2.0) * (xx.astype(np.float32) + np.float32(0.5))) / np.float32(w) - np.float32(1.0)
    pny = np.float32(1.0) - (np.float32(2.0) * (yy.astype(np.float32) + np.float32(0.5))) / np.float32(h)
    for i in range(n):
        p = proj[i]
        if p[0] == 0:
            continue
        single = np.zeros_like(proj)
        single[i] = p
        counts = oracle.tile_counts(single, w, h)
        det = p[3] * p[6] - p[5] * p[4]
        inv = np.array([[p[6] / det, -p[5] / det], [-p[4] / det, p[3] / det]], np.float64)
        dx, dy = pnx - p[1], pny - p[2]
        r2 = (inv[0, 0] * dx + inv[0, 1] * dy) ** 2 + (inv[1, 0] * dx + inv[1, 1] * dy) ** 2
        cut = oracle.coverage_cut(float(p[10]))
        covered = r2 <= cut * (1 + 1e-5) + 1e-7
        tiles_hit = np.zeros(counts.shape, bool)
        ys, xs = np.nonzero(covered)
        tiles_hit[ys // 16, xs // 16] = True
        assert np.all(counts[tiles_hit] == 1), i


def test_blend_evals_single_splat_counts_covered_pixels(oracle):
    # One splat: every covered pixel is evaluated once and none stops early,
    # so E is the number of pixels the oracle's own raster covers.
    s = oracle.Scene([[0.1, -0.05, 2.0]], [1.0], [[-2.4, -2.9, -2.6]], [[0.2, 0.1, 0.0, 1.0]], [[0.3, -0.2, 0.1]])
    w, h = 120, 90
    img, _ = oracle.render(s, default_cam(), w, h)
    assert oracle.blend_evals(s, default_cam(), w, h) == int((img[..., 3] > 0).sum()) > 100


def test_blend_evals_opaque_stack_stops_early(oracle):
    # Eight near-opaque splats (alpha 0.9995) stacked in depth: near the centre
    # (alpha*g ~ 0.95) T falls below 1e-4 after 3-4 of them, towards the rim
    # all eight are evaluated; brute force in float64 per pixel.
    k = 8
    pos = [[0.0, 0.0, 2.0 + 0.1 * i] for i in range(k)]
    logit = math.log(0.9995 / 0.0005)
    s = oracle.Scene(pos, [logit] * k, [[-2.5] * 3] * k, [[0, 0, 0, 1]] * k, [[0.2, 0.2, 0.2]] * k)
    w = 64
    e = oracle.blend_evals(s, default_cam(), w, w)
    proj = oracle.project(s, default_cam(), w, w)
    yy, xx = np.mgrid[0:w, 0:w]
    pnx = 2 * (xx + 0.5) / w - 1
    pny = 1 - 2 * (yy + 0.5) / w
    T = np.ones((w, w))
    count = np.zeros((w, w), np.int64)
    for i in range(k):  # nearest first: index 0 is nearest
        p = proj[i]
        inv = np.linalg.inv(np.array([[p[3], p[5]], [p[4], p[6]]], np.float64))
        dx, dy = pnx - p[1], pny - p[2]
        r2 = (inv[0, 0] * dx + inv[0, 1] * dy) ** 2 + (inv[1, 0] * dx + inv[1, 1] * dy) ** 2
        m = (r2 <= oracle.coverage_cut(float(p[10]))) & (T >= 1e-4)
        count += m
        T = np.where(m, T * (1 - p[10] * np.exp(-4.5 * r2)), T)
    assert abs(e - int(count.sum())) <= 4, (e, int(count.sum()))
    assert count.max() == k and 2 <= count[w // 2, w // 2] < 5


def test_blend_evals_brute_force_random(oracle):
    rng = np.random.default_rng(19)
    n = 60
    pos = np.c_[rng.normal(0, 0.25, (n, 2)), rng.uniform(1.5, 3.0, n)]
    s = oracle.Scene(pos, rng.normal(1.0, 1.5, n), rng.normal(-2.6, 0.4, (n, 3)),
                     rng.standard_normal((n, 4)), rng.normal(0, 0.3, (n, 3)))
    w, h = 72, 56
    e = oracle.blend_evals(s, default_cam(), w, h)
    proj = oracle.project(s, default_cam(), w, h)
    keys, idx = oracle.preprocess(s, default_cam())
    order = oracle.sort_values_by_keys(keys, idx)[::-1]
    yy, xx = np.mgrid[0:h, 0:w]
    pnx = 2 * (xx + 0.5) / w - 1
    pny = 1 - 2 * (yy + 0.5) / h
    T = np.ones((h, w))
    total = 0
    for i in order:
        p = proj[i]
        if p[0] == 0:
            continue
        inv = np.linalg.inv(np.array([[p[3], p[5]], [p[4], p[6]]], np.float64))
        dx, dy = pnx - p[1], pny - p[2]
        r2 = (inv[0, 0] * dx + inv[0, 1] * dy) ** 2 + (inv[1, 0] * dx + inv[1, 1] * dy) ** 2
        m = (r2 <= min(oracle.coverage_cut(float(p[10])), 1.0)) & (T >= 1e-4)
        total += int(m.sum())
        T = np.where(m, T * (1 - p[10] * np.exp(-4.5 * r2)), T)
    assert total > 1000
    assert abs(e - total) <= max(3, total // 2000), (e, total)
