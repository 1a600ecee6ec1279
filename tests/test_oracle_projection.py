"""Pins the oracle's projection (project_splat, splat_surface_direct.wgsl:131-217;
CPU form render-wgpu/src/lib.rs:1456-1506, 1589-1634) off the optical axis and
at every clamp, against an independent float64 model.

The model knows only the mathematics the reference states in prose:
  * a perspective camera, NDC = (f x / (z aspect), f y / z) with f = 1/tan(fov/2);
  * the EWA image covariance J Sigma_cam J^T + a 0.3-pixel blur, J the Jacobian of
    that map taken by central finite differences in float64 (no closed form), at
    the view-space point with x/z, y/z clamped to 1.3 tan(fov/2) (times aspect in x);
  * axes = 3 sigma of the image covariance's eigen-decomposition (numpy eigh),
    radii clamped to [1e-4, 2] and the major to 64 times the minor; when the
    covariance's off-diagonal term is at most 1e-8 the axes are taken along x and
    y (the larger variance's axis first), as the reference states;
  * rotations from scipy's quaternion code (not the oracle's quat_to_mat3).
It is the mathematics, not the oracle's float32 expression order, so the checks
hold to a few float32 ulps of the largest term. Each check fails when the
corresponding mistake is planted in the oracle (tests/test_oracle_mutations.py).
"""

import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from helpers import DEFAULT_FOV

FOV = float(DEFAULT_FOV)


def cam_dict(position, quat_xyzw, fov=FOV):
    q = np.asarray(quat_xyzw, np.float64)
    q = (q / np.linalg.norm(q)).astype(np.float32)
    return dict(position=[float(np.float32(v)) for v in position], rotation_xyzw=[float(v) for v in q], fov=fov,
                near=0.01, far=1000.0)


# A camera turned away from every axis, so that view rows, x and y all matter.
CAM = cam_dict([0.3, -0.4, -2.5], [0.12, -0.31, 0.07, 0.94])
IDENT = cam_dict([0.0, 0.0, 0.0], [0.0, 0.0, 0.0, 1.0])


def view_matrix(cam):
    """World-to-camera rotation: the inverse of the camera's orientation (scipy)."""
    return Rotation.from_quat(np.asarray(cam["rotation_xyzw"], np.float64)).as_matrix().T


def world_from_cam(cam, pc):
    """World position (float32) of view-space point pc."""
    v = view_matrix(cam)
    return (np.asarray(cam["position"], np.float64) + v.T @ np.asarray(pc, np.float64)).astype(np.float32)


def world_cov(log_scale, quat_xyzw):
    r = Rotation.from_quat(np.asarray(quat_xyzw, np.float64)).as_matrix()
    s = np.maximum(np.exp(np.asarray(log_scale, np.float64)), 1e-6)
    return r @ np.diag(s * s) @ r.T


def model(cam, pos, log_scale, quat, w, h, clamp=True):
    """Float64 statement of the projection: (ndc centre, image covariance, view-space point)."""
    v = view_matrix(cam)
    pc = v @ (np.asarray(pos, np.float64) - np.asarray(cam["position"], np.float64))
    cov_c = v @ world_cov(log_scale, quat) @ v.T
    t = math.tan(cam["fov"] / 2.0)
    f, aspect = 1.0 / t, w / h

    def ndc(q):
        return np.array([f * q[0] / (q[2] * aspect), f * q[1] / q[2]])

    z = pc[2]
    q = pc.copy()
    if clamp:
        q[0] = np.clip(pc[0] / z, -1.3 * t * aspect, 1.3 * t * aspect) * z
        q[1] = np.clip(pc[1] / z, -1.3 * t, 1.3 * t) * z
    jac = np.zeros((2, 3))
    for k in range(3):
        e = np.zeros(3)
        e[k] = 1e-6 * max(1.0, abs(q[k]))
        jac[:, k] = (ndc(q + e) - ndc(q - e)) / (2.0 * e[k])
    cov2 = jac @ cov_c @ jac.T + np.diag([(0.6 / w) ** 2, (0.6 / h) ** 2])
    return ndc(pc), cov2, pc


def model_frame(cov2):
    """Eigenvalues (ascending) and unit eigenvectors (columns) of the image covariance,
    axis-aligned when |off-diagonal| <= 1e-8 (the major along x unless var_x < var_y)."""
    lam, vec = np.linalg.eigh(cov2)
    if abs(cov2[0, 1]) <= 1e-8:
        vec = np.array([[0.0, 1.0], [1.0, 0.0]]) if cov2[0, 0] >= cov2[1, 1] else np.eye(2)
    return lam, vec


def model_axes(cov2):
    """3-sigma axes with the stated clamps: (u, v) as 2-vectors."""
    lam, vec = model_frame(cov2)
    r_min = min(max(3.0 * math.sqrt(max(lam[0], 1e-10)), 1e-4), 2.0)
    r_maj = min(max(3.0 * math.sqrt(max(lam[1], 1e-10)), 1e-4), 2.0)
    r_maj = min(r_maj, 64.0 * r_min)
    return vec[:, 1] * r_maj, vec[:, 0] * r_min


def project_one(oracle, cam, pos, log_scale, quat, w, h, opacity=2.0):
    s = oracle.Scene([pos], [opacity], [log_scale], [quat], [[0.1, 0.2, 0.3]])
    return oracle.project(s, cam, w, h)[0]


def axes_outer(p):
    """A A^T of the oracle's axes, A = [axis_u axis_v]."""
    u, v = p[3:5].astype(np.float64), p[5:7].astype(np.float64)
    return np.outer(u, u) + np.outer(v, v)


def assert_outer_matches(p, cov2, label, rel=2e-5):
    lam, vec = model_frame(cov2)
    got, want = axes_outer(p), 9.0 * (vec * lam) @ vec.T  # = 9 cov2 off the axis-aligned branch
    tol = rel * np.abs(want).max()
    assert np.abs(got - want).max() <= tol, f"{label}: A A^T {got.ravel()} vs 9 cov2 {want.ravel()}"


# -------------------------------------------------------- off-axis EWA

OFF_AXIS = [
    # view-space position, log-scales, splat quaternion (xyzw)
    ((0.6, -0.4, 2.0), (-2.5, -2.5, -2.5), (0.0, 0.0, 0.0, 1.0)),   # isotropic, off-axis in x and y
    ((-0.9, 0.5, 3.0), (-1.8, -3.0, -2.2), (0.3, -0.5, 0.2, 0.8)),  # anisotropic, rotated (xz, yz != 0)
    ((0.3, 0.7, 1.6), (-2.0, -2.9, -1.6), (-0.6, 0.1, 0.4, 0.7)),
    ((1.1, -0.6, 2.4), (-1.5, -3.5, -3.5), (0.9, 0.2, -0.3, 0.2)),  # elongated
    ((0.45, 0.35, 3.0), (-7.0, -7.0, -7.0), (0.0, 0.0, 0.0, 1.0)),  # tiny: the blur term dominates
]


@pytest.mark.parametrize("case", range(len(OFF_AXIS)))
@pytest.mark.parametrize("size", [(1280, 720), (480, 640)])
def test_offaxis_axes_match_finite_difference_jacobian(oracle, case, size):
    pc, ls, q = OFF_AXIS[case]
    w, h = size
    pos = world_from_cam(CAM, pc)
    p = project_one(oracle, CAM, pos, ls, q, w, h)
    center, cov2, _ = model(CAM, pos.astype(np.float64), np.float32(ls), np.float32(q), w, h)
    assert p[0] == 1, "splat must be drawn"
    ru, rv = np.hypot(p[3], p[4]), np.hypot(p[5], p[6])
    assert 1.01e-4 < rv <= ru < 1.99 and ru < 63.0 * rv, "no clamp may be active for this check"
    assert np.allclose(p[1:3], center, rtol=0, atol=1e-5 * max(1.0, np.abs(center).max())), (p[1:3], center)
    assert_outer_matches(p, cov2, f"case {case} {w}x{h}")
    assert abs(p[3] * p[5] + p[4] * p[6]) <= 1e-5 * ru * ru  # axes orthogonal


def test_reference_offaxis_covariance_kat(oracle):
    # render-wgpu/src/lib.rs:4442-4461: camera covariance diag(0.020, 0.005, 0.002) at
    # (0.2, -0.1, 2.0), default camera and 1280x720 config: both axes non-zero, |u| > |v|.
    diag = np.array([0.020, 0.005, 0.002])
    ls = (0.5 * np.log(diag)).astype(np.float32)
    p = project_one(oracle, IDENT, [0.2, -0.1, 2.0], ls, [0, 0, 0, 1], 1280, 720)
    lu, lv = np.hypot(p[3], p[4]), np.hypot(p[5], p[6])
    assert p[0] == 1 and lu > 0 and lv > 0 and lu > lv
    _, cov2, _ = model(IDENT, [0.2, -0.1, 2.0], ls, [0, 0, 0, 1], 1280, 720)
    assert_outer_matches(p, cov2, "reference covariance KAT")


# ------------------------------------------------ the 1.3 tan(fov/2) clamp

def _clamp_case(oracle, xz_frac, yz_frac, log_scale, w=1280, h=720):
    """Splat at x/z = xz_frac * lim_x, y/z = yz_frac * lim_y (z = 2) of the identity camera."""
    t = math.tan(FOV / 2.0)
    lim_x, lim_y = 1.3 * t * (w / h), 1.3 * t
    z = 2.0
    pos = np.array([xz_frac * lim_x * z, yz_frac * lim_y * z, z], np.float32)
    ls = np.float32([log_scale] * 3)
    p = project_one(oracle, IDENT, pos, ls, [0, 0, 0, 1], w, h)
    _, cov_c, _ = model(IDENT, pos.astype(np.float64), ls, [0, 0, 0, 1], w, h, clamp=True)
    _, cov_u, _ = model(IDENT, pos.astype(np.float64), ls, [0, 0, 0, 1], w, h, clamp=False)
    return p, cov_c, cov_u


@pytest.mark.parametrize("xf,yf,ls", [(1.15, 0.3, -0.9), (0.4, 1.2, -0.9), (-1.25, -1.1, -0.7)])
def test_jacobian_clamp_at_frustum_edge(oracle, xf, yf, ls):
    # Beyond the clamp the Jacobian is evaluated at the clamped point (and the
    # centre is not clamped); the unclamped Jacobian gives a different ellipse.
    p, cov_c, cov_u = _clamp_case(oracle, xf, yf, ls)
    assert p[0] == 1
    ru, rv = np.hypot(p[3], p[4]), np.hypot(p[5], p[6])
    assert rv <= ru < 1.99
    assert_outer_matches(p, cov_c, f"clamped ({xf}, {yf})")
    gap = np.abs(axes_outer(p) - 9.0 * cov_u).max() / np.abs(9.0 * cov_u).max()
    assert gap > 1e-2, "the unclamped Jacobian must differ here"


def test_jacobian_unclamped_inside_aspect_widened_limit(oracle):
    # x/z between 1.3 tan(fov/2) and 1.3 tan(fov/2) * aspect: inside the x limit,
    # so the Jacobian is the unclamped one (the limit is widened by the aspect).
    t = math.tan(FOV / 2.0)
    w, h = 1280, 720
    frac = 1.1 * (1.3 * t) / (1.3 * t * w / h)  # x/z = 1.1 * 1.3 tan
    p, cov_c, cov_u = _clamp_case(oracle, frac, 0.2, -2.3, w, h)
    assert p[0] == 1
    assert np.array_equal(cov_c, cov_u)
    assert_outer_matches(p, cov_u, "inside the widened x limit")


# ------------------------------------------------------- radius clamps

def test_radius_upper_clamp(oracle):
    # A splat far larger than the view: both radii are clamped to exactly 2.
    p = project_one(oracle, IDENT, [0.1, 0.05, 2.0], [1.0, 1.2, 0.8], [0.2, 0.1, 0.0, 0.97], 1280, 720)
    assert p[0] == 1
    assert abs(np.hypot(p[3], p[4]) - 2.0) < 1e-6 and abs(np.hypot(p[5], p[6]) - 2.0) < 1e-6


def test_radius_lower_clamp(oracle):
    # At 40000x40000 a sub-pixel splat's 3 sigma (blur included) is below 1e-4:
    # both radii are clamped to 1e-4.
    w = h = 40000
    pos, ls = [0.01, -0.01, 3.0], [-14.0, -14.0, -14.0]
    p = project_one(oracle, IDENT, pos, ls, [0, 0, 0, 1], w, h)
    _, cov2, _ = model(IDENT, pos, np.float32(ls), [0, 0, 0, 1], w, h)
    assert 3.0 * math.sqrt(np.linalg.eigvalsh(cov2).max()) < 0.6e-4  # unclamped radius
    assert p[0] == 1
    assert abs(np.hypot(p[3], p[4]) - 1e-4) < 1e-10 and abs(np.hypot(p[5], p[6]) - 1e-4) < 1e-10


def test_anisotropy_cap(oracle):
    # A 10^4:1 needle along the image x axis: the major radius is capped at 64 x the minor.
    w, h = 1280, 720
    pos, ls = [0.05, -0.02, 4.0], np.float32([0.0, math.log(1e-4), math.log(1e-4)])
    p = project_one(oracle, IDENT, pos, ls, [0, 0, 0, 1], w, h)
    _, cov2, _ = model(IDENT, pos, ls, [0, 0, 0, 1], w, h)
    lam = np.linalg.eigvalsh(cov2)
    assert 3.0 * math.sqrt(lam[1]) > 200.0 * 3.0 * math.sqrt(lam[0])  # far beyond 64:1 uncapped
    assert p[0] == 1
    ru, rv = np.hypot(p[3], p[4]), np.hypot(p[5], p[6])
    assert abs(ru / rv - 64.0) < 1e-4
    # the minor eigenvalue is the float32 difference (a + c)/2 - |a - c|/2 of terms
    # ~10^5 times larger, so the minor radius holds to ~1e-3 only
    assert abs(rv - 3.0 * math.sqrt(lam[0])) < 5e-3 * rv
    u, v = model_axes(cov2)
    assert abs(abs(p[3]) - abs(u[0])) < 5e-3 * ru and abs(abs(p[6]) - abs(v[1])) < 5e-3 * rv


# --------------------------------------------------- NDC-box cull edges

@pytest.mark.parametrize("direction", [(1, 0), (-1, 0), (0, 1), (0, -1)])
def test_ndc_box_cull_edges(oracle, direction):
    # A rotated ellipse slides out of the view along `direction`; the last drawn and
    # first culled float32 positions straddle the point where the box
    # centre +- (|u_x| + |v_x|, |u_y| + |v_y|) of the model's axes leaves [-1, 1]^2.
    w, h = 1280, 720
    z, ls, q = 2.5, np.float32([-2.2, -3.1, -3.1]), np.float32([0.0, 0.0, 0.38, 0.92])
    dx, dy = direction
    t = math.tan(FOV / 2.0)
    span = 1.3 * t * z * (w / h if dx else 1.0)

    def pos_at(s):
        return np.array([dx * s + (0.0 if dx else 0.07), dy * s + (0.0 if dy else -0.05), z], np.float32)

    def drawn(s):
        return project_one(oracle, IDENT, pos_at(s), ls, q, w, h)[0] == 1

    lo, hi = 0.0, span
    assert drawn(lo) and not drawn(hi)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if pos_at(mid)[0 if dx else 1] in (pos_at(lo)[0 if dx else 1], pos_at(hi)[0 if dx else 1]):
            break
        if drawn(mid):
            lo = mid
        else:
            hi = mid
    axis = 0 if dx else 1
    sign = dx or dy
    for s, inside in ((lo, True), (hi, False)):
        center, cov2, _ = model(IDENT, pos_at(s).astype(np.float64), ls, q, w, h)
        u, v = model_axes(cov2)
        ext = abs(u[axis]) + abs(v[axis])
        edge = sign * center[axis] - ext  # > 1: the box misses the view
        assert abs(edge - 1.0) < 2e-5, (s, edge)
        assert min(abs(u[axis]), abs(v[axis])) > 1e-2, "both axes must reach across the edge"


# ------------------------------------------- tile rectangles, both sides

def _ellipse_extent_px(p, cut, w, h):
    """Pixel-index extents of {centre + A l : |l|^2 <= cut}, sampled densely on its boundary."""
    th = np.linspace(0.0, 2.0 * math.pi, 8192, endpoint=False)
    lx, ly = math.sqrt(cut) * np.cos(th), math.sqrt(cut) * np.sin(th)
    ndc_x = p[1] + p[3] * lx + p[5] * ly
    ndc_y = p[2] + p[4] * lx + p[6] * ly
    px = (ndc_x + 1.0) * w / 2.0 - 0.5
    py = (1.0 - ndc_y) * h / 2.0 - 0.5
    return px.min(), px.max(), py.min(), py.max()


def test_tile_rect_two_sided(oracle):
    # At tile size 1 the counts are the splat's pixel rectangle. It must hold every
    # covered pixel (brute force in float64), and each edge must be the extent of
    # the coverage ellipse (its boundary sampled densely) plus one pixel, rounded
    # inwards to pixel centres, within 1e-3 px; at 16 px the counted tiles are
    # exactly the tiles of that rectangle.
    rng = np.random.default_rng(23)
    n = 160
    pos = np.c_[rng.normal(0, 0.6, (n, 2)), rng.uniform(1.2, 4.0, n)].astype(np.float32)
    s = oracle.Scene(pos, rng.normal(0.5, 1.5, n), rng.normal(-2.8, 0.7, (n, 3)),
                     rng.standard_normal((n, 4)), rng.normal(0, 0.3, (n, 3)))
    w, h = 192, 128
    proj = oracle.project(s, IDENT, w, h)
    yy, xx = np.mgrid[0:h, 0:w]
    pnx = (np.float32(2.0) * (xx.astype(np.float32) + np.float32(0.5))) / np.float32(w) - np.float32(1.0)
    pny = np.float32(1.0) - (np.float32(2.0) * (yy.astype(np.float32) + np.float32(0.5))) / np.float32(h)
    checked = 0
    for i in range(n):
        p = proj[i]
        if p[0] == 0:
            continue
        single = np.zeros_like(proj)
        single[i] = p
        px = oracle.tile_counts(single, w, h, tile=1)
        tiles = oracle.tile_counts(single, w, h, tile=16)
        cut = min(oracle.coverage_cut(float(p[10])), 1.0)
        ys, xs = np.nonzero(px)
        if cut < 0:
            assert xs.size == 0
            continue
        inv = np.linalg.inv(np.array([[p[3], p[5]], [p[4], p[6]]], np.float64))
        dx, dy = pnx - p[1], pny - p[2]
        r2 = (inv[0, 0] * dx + inv[0, 1] * dy) ** 2 + (inv[1, 0] * dx + inv[1, 1] * dy) ** 2
        cov = r2 <= cut * (1 + 1e-5) + 1e-7
        assert np.all(px[cov] == 1), f"splat {i}: a covered pixel outside its rectangle"
        if xs.size == 0:
            continue
        x0e, x1e, y0e, y1e = _ellipse_extent_px(p, cut, w, h)
        x0, x1, y0, y1 = xs.min(), xs.max(), ys.min(), ys.max()
        assert px[y0:y1 + 1, x0:x1 + 1].all() and px.sum() == (x1 - x0 + 1) * (y1 - y0 + 1), \
            f"splat {i}: not a rectangle"
        eps = 1e-3
        assert max(0, math.ceil(x0e - 1.0 - eps)) <= x0 <= max(0, math.ceil(x0e - 1.0 + eps)), (i, x0, x0e)
        assert min(w - 1, math.floor(x1e + 1.0 - eps)) <= x1 <= min(w - 1, math.floor(x1e + 1.0 + eps)), (i, x1, x1e)
        assert max(0, math.ceil(y0e - 1.0 - eps)) <= y0 <= max(0, math.ceil(y0e - 1.0 + eps)), (i, y0, y0e)
        assert min(h - 1, math.floor(y1e + 1.0 - eps)) <= y1 <= min(h - 1, math.floor(y1e + 1.0 + eps)), (i, y1, y1e)
        want = np.zeros_like(tiles)
        want[y0 // 16:y1 // 16 + 1, x0 // 16:x1 // 16 + 1] = 1
        assert np.array_equal(tiles, want), f"splat {i}: 16 px tiles are not those of the rectangle"
        checked += 1
    assert checked > 100
