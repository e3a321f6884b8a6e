"""Seeded synthetic test images (numpy): smooth value noise + Gaussian blobs +
rectangles, and simple warps. Small and fast; the bench uses the library's
own C++ pair generator."""
from __future__ import annotations

import numpy as np


def value_noise(rng, h, w, cell, amp):
    gh, gw = h // cell + 2, w // cell + 2
    lat = rng.uniform(-amp, amp, (gh, gw))
    yy, xx = np.mgrid[0:h, 0:w] / cell
    y0, x0 = np.floor(yy).astype(int), np.floor(xx).astype(int)
    fy, fx = yy - y0, xx - x0
    return (lat[y0, x0] * (1 - fx) * (1 - fy) + lat[y0, x0 + 1] * fx * (1 - fy) +
            lat[y0 + 1, x0] * (1 - fx) * fy + lat[y0 + 1, x0 + 1] * fx * fy)


def texture_rgb(seed: int, h: int, w: int, blobs: int = 60, rects: int = 15) -> np.ndarray:
    rng = np.random.default_rng(seed)
    img = np.full((h, w, 3), 128.0) + value_noise(rng, h, w, 16, 20)[..., None]
    for _ in range(rects):
        y0, x0 = rng.integers(0, h), rng.integers(0, w)
        img[y0:y0 + rng.integers(4, max(5, h // 4)), x0:x0 + rng.integers(4, max(5, w // 4))] = rng.uniform(0, 255, 3)
    for _ in range(blobs):
        cy, cx, s = rng.uniform(0, h), rng.uniform(0, w), rng.uniform(3, 20)
        r = int(4 * s) + 1
        y0, y1 = max(0, int(cy) - r), min(h, int(cy) + r)
        x0, x1 = max(0, int(cx) - r), min(w, int(cx) + r)
        if y0 >= y1 or x0 >= x1:
            continue
        yy, xx = np.mgrid[y0:y1, x0:x1]
        g = np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * s * s))
        img[y0:y1, x0:x1] += g[..., None] * rng.uniform(-120, 120, 3)
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def warp_rgb(src: np.ndarray, H: np.ndarray, h: int, w: int, seed: int = 0, noise: float = 2.0) -> np.ndarray:
    """B(x) = bilinear src(H^-1 x) (H maps src -> B), plus Gaussian noise."""
    rng = np.random.default_rng(seed)
    Hi = np.linalg.inv(H)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    den = Hi[2, 0] * xx + Hi[2, 1] * yy + Hi[2, 2]
    sx = (Hi[0, 0] * xx + Hi[0, 1] * yy + Hi[0, 2]) / den
    sy = (Hi[1, 0] * xx + Hi[1, 1] * yy + Hi[1, 2]) / den
    sh, sw = src.shape[:2]
    sx = np.clip(sx, 0, sw - 1.001)
    sy = np.clip(sy, 0, sh - 1.001)
    x0, y0 = np.floor(sx).astype(int), np.floor(sy).astype(int)
    fx, fy = (sx - x0)[..., None], (sy - y0)[..., None]
    s = src.astype(np.float64)
    out = (s[y0, x0] * (1 - fx) * (1 - fy) + s[y0, x0 + 1] * fx * (1 - fy) +
           s[y0 + 1, x0] * (1 - fx) * fy + s[y0 + 1, x0 + 1] * fx * fy)
    out += rng.normal(0, noise, out.shape)
    return np.clip(np.rint(out), 0, 255).astype(np.uint8)


def gray_of(rgb: np.ndarray) -> np.ndarray:
    r, g, b = (rgb[..., i].astype(np.int64) for i in range(3))
    return ((9798 * r + 19235 * g + 3735 * b + 16384) >> 15).astype(np.uint8)
