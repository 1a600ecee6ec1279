"""Deterministic synthetic scenes and cameras of the benchmark workloads.

Seeds are 0x6A5F0000 + workload index; scenes inside a workload are sub-seeded
by their index. All attributes are float32 in the reference SceneBuffers
layout (crates/gsplat-core/src/lib.rs:201-213): positions xyz, opacity logit,
log-scales, xyzw quaternions, DC colour and channel-major SH rest.

Workloads:
  plain-640x480-16k-sh0     16,384 splats, SH0, one 640x480 view
  plain-1080p-300k-sh3      300,000 splats, SH3, one 1920x1080 view
  plain-4k-3m-sh3           3,000,000 splats, SH3 heavy-tailed scales, 3840x2160
  batched-64v-8s            8 scenes (50k..1M), 64 views of mixed sizes
  served-16s-10m            16 scenes totalling 10M splats, Poisson requests
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

SEED_BASE = 0x6A5F0000
RESOLUTIONS = [(320, 240), (640, 480), (1280, 720), (1920, 1080)]


@dataclass
class Scene:
    positions: np.ndarray
    opacity: np.ndarray
    scale_xyz: np.ndarray
    rotation_xyzw: np.ndarray
    color_dc: np.ndarray
    sh_degree: int
    sh_rest: np.ndarray | None

    @property
    def count(self) -> int:
        return int(self.positions.shape[0])

    def subset(self, n: int) -> "Scene":
        return Scene(self.positions[:n].copy(), self.opacity[:n].copy(), self.scale_xyz[:n].copy(),
                     self.rotation_xyzw[:n].copy(), self.color_dc[:n].copy(), self.sh_degree,
                     None if self.sh_rest is None else self.sh_rest[:n].copy())


@dataclass
class View:
    position: tuple
    target: tuple
    width: int
    height: int
    scene: int = 0


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _quats(rng, n):
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return _f32(q)


def gaussian_scene(rng, n, sh_degree, means, log_scale_mean=-4.0, log_scale_std=0.5, heavy_tail=0.0):
    """Splats with the attribute laws shared by all workloads."""
    scales = rng.normal(log_scale_mean, log_scale_std, (n, 3))
    if heavy_tail > 0.0:
        heavy = rng.random(n) < heavy_tail
        scales[heavy] = -2.0
    opacity = rng.normal(0.0, 1.5, n)
    dc = rng.normal(0.0, 0.3, (n, 3))
    rest = None
    if sh_degree > 0:
        k = 3 * ((sh_degree + 1) ** 2 - 1)
        rest = _f32(rng.normal(0.0, 0.05, (n, k)))
    return Scene(_f32(means), _f32(opacity), _f32(scales), _quats(rng, n), _f32(dc), sh_degree, rest)


def mixture_means(rng, n, radius=1.5, sigma=0.5, components=3):
    """Equal-weight Gaussian mixture with centres uniform on a sphere of `radius`."""
    c = rng.standard_normal((components, 3))
    c = radius * c / np.linalg.norm(c, axis=1, keepdims=True)
    which = rng.integers(0, components, n)
    return c[which] + rng.normal(0.0, sigma, (n, 3))


def sphere_point(rng, radius):
    v = rng.standard_normal(3)
    return tuple(float(x) for x in radius * v / np.linalg.norm(v))


def plain_small(seed_offset=0):
    """16,384 splats SH0, 640x480, camera at (0,0,-3) looking at the origin."""
    rng = np.random.default_rng(SEED_BASE + 0 + seed_offset)
    n = 16384
    scene = gaussian_scene(rng, n, 0, rng.normal(0.0, 0.5, (n, 3)))
    return scene, View((0.0, 0.0, -3.0), (0.0, 0.0, 0.0), 640, 480)


def plain_1080p(n=300_000, seed_offset=0):
    """300,000 splats SH3, 1920x1080, camera on a sphere of radius 4."""
    rng = np.random.default_rng(SEED_BASE + 1 + seed_offset)
    scene = gaussian_scene(rng, n, 3, mixture_means(rng, n))
    return scene, View(sphere_point(rng, 4.0), (0.0, 0.0, 0.0), 1920, 1080)


def plain_4k(n=3_000_000, seed_offset=0):
    """3,000,000 splats SH3, log-scales N(-5, 0.7^2) with 10% at -2, 3840x2160."""
    rng = np.random.default_rng(SEED_BASE + 2 + seed_offset)
    scene = gaussian_scene(rng, n, 3, mixture_means(rng, n), -5.0, 0.7, heavy_tail=0.1)
    return scene, View(sphere_point(rng, 4.0), (0.0, 0.0, 0.0), 3840, 2160)


BATCH_SCENE_SIZES = [50_000, 100_000, 200_000, 300_000, 500_000, 700_000, 850_000, 1_000_000]


def batched(views=64, sizes=None):
    """8 resident scenes drawn as the 1080p workload, `views` views; view v renders scene v mod 8."""
    sizes = sizes or BATCH_SCENE_SIZES
    rng = np.random.default_rng(SEED_BASE + 3)
    scenes = []
    for s, n in enumerate(sizes):
        r = np.random.default_rng(SEED_BASE + 3 + 1000 * (s + 1))
        scenes.append(gaussian_scene(r, n, 3, mixture_means(r, n)))
    vs = []
    for v in range(views):
        w, h = RESOLUTIONS[int(rng.integers(0, len(RESOLUTIONS)))]
        vs.append(View(sphere_point(rng, 4.0), (0.0, 0.0, 0.0), w, h, v % len(sizes)))
    return scenes, vs


def served_sizes(total=10_000_000, count=16, lo=100_000, hi=2_000_000, seed=SEED_BASE + 4):
    rng = np.random.default_rng(seed)
    s = np.exp(rng.uniform(math.log(lo), math.log(hi), count))
    s = np.maximum((s * total / s.sum()).astype(np.int64), 1)
    s[-1] += total - s.sum()
    return [int(v) for v in s]


def served(total=10_000_000, count=16, requests=512, rate=5000.0):
    """16 scenes with log-uniform sizes totalling `total`; Poisson arrivals of single-view requests."""
    sizes = served_sizes(total, count)
    scenes = []
    for s, n in enumerate(sizes):
        r = np.random.default_rng(SEED_BASE + 4 + 1000 * (s + 1))
        scenes.append(gaussian_scene(r, n, 3, mixture_means(r, n)))
    rng = np.random.default_rng(SEED_BASE + 4)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, requests))
    reqs = []
    for i in range(requests):
        w, h = RESOLUTIONS[int(rng.integers(0, len(RESOLUTIONS)))]
        radius = float(rng.uniform(3.0, 6.0))
        reqs.append((float(arrivals[i]), View(sphere_point(rng, radius), (0.0, 0.0, 0.0), w, h,
                                              int(rng.integers(0, count)))))
    return scenes, reqs


def tiny_scene(n, seed, sh_degree=0, spread=0.5, log_scale=-2.5):
    """Small scenes for parity tests."""
    rng = np.random.default_rng(seed)
    return gaussian_scene(rng, n, sh_degree, rng.normal(0.0, spread, (n, 3)), log_scale, 0.5)
