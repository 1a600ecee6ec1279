"""Image comparison metrics of the reference's parity gates."""

import numpy as np


def rgba_parity_metrics(a8, b8):
    """rgba_image_parity_metrics (render-wgpu/src/lib.rs:3491-3524) of two RGBA8 images: the mean
    |difference| over the RGB channels, the fraction of pixels with any RGB channel off by more than
    3/255 (pixels, not channels: lib.rs:3539-3545), and the largest |difference|, in units of 1."""
    a = np.asarray(a8).reshape(-1, 4)[:, :3].astype(np.float64) / 255.0
    b = np.asarray(b8).reshape(-1, 4)[:, :3].astype(np.float64) / 255.0
    if a.shape != b.shape:
        raise ValueError("images differ in size")
    if not a.size:
        return 0.0, 0.0, 0.0
    d = np.abs(a - b)
    return float(d.mean()), float((d > 3.0 / 255.0).any(axis=1).mean()), float(d.max())


def passes_stale_order_gate(a8, b8):
    """The reference's bar for an image drawn from a stale order or a lossy scene codec
    (assert_image_parity, lib.rs:3526-3537): mean |delta RGB| <= 1/255 and at most 0.1% of pixels
    off by more than 3/255."""
    mean, frac, _ = rgba_parity_metrics(a8, b8)
    return mean <= 1.0 / 255.0 and frac <= 0.001
