"""Writes the golden vectors under tests/golden/ from the reference's own NumPy
oracles. Run in a checkout where the reference repository is available:

    python tests/golden/make_golden.py --reference /root/reference [--only s6]

It imports two numpy-only modules of the reference,
tools/s2_s3/reference.py (single-level ORB: detect + describe) and
tools/s6_reference/reference.py (FP64 warp / feather / multiband render),
runs them on small inputs built here, and stores inputs and outputs as .npz
data. Nothing of the reference is stored except these input/output vectors.
The committed fixtures are what the tests read; this script is only rerun to
regenerate them.
"""
from __future__ import annotations

import argparse
import importlib
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def _smooth_noise(rng, h, w, cell):
    gh, gw = h // cell + 2, w // cell + 2
    lattice = rng.uniform(0, 255, (gh, gw))
    yy, xx = np.mgrid[0:h, 0:w] / cell
    y0, x0 = np.floor(yy).astype(int), np.floor(xx).astype(int)
    fy, fx = yy - y0, xx - x0
    v = (lattice[y0, x0] * (1 - fx) * (1 - fy) + lattice[y0, x0 + 1] * fx * (1 - fy) +
         lattice[y0 + 1, x0] * (1 - fx) * fy + lattice[y0 + 1, x0 + 1] * fx * fy)
    return v


def gray_cases():
    """Small gray images exercising FAST/NMS/ties/borders."""
    rng = np.random.default_rng(20261022)
    cases = {}
    # Textured: smooth noise plus blobs and rectangles.
    h, w = 120, 160
    img = _smooth_noise(rng, h, w, 9)
    for _ in range(25):
        cy, cx, s = rng.uniform(0, h), rng.uniform(0, w), rng.uniform(2, 8)
        yy, xx = np.mgrid[0:h, 0:w]
        img += rng.uniform(-90, 90) * np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * s * s))
    for _ in range(8):
        y0, x0 = rng.integers(0, h - 10), rng.integers(0, w - 10)
        img[y0:y0 + rng.integers(4, 30), x0:x0 + rng.integers(4, 30)] = rng.uniform(0, 255)
    cases["textured_160x120"] = np.clip(np.rint(img), 0, 255).astype(np.uint8)
    # Checkerboard with 5-px cells: many equal FAST scores (tie handling).
    yy, xx = np.mgrid[0:96, 0:112]
    cases["checker_112x96"] = (((yy // 5 + xx // 5) % 2) * 200 + 20).astype(np.uint8)
    # Pure noise: dense corners, boundary-score ties in the 2K pre-selection.
    cases["noise_100x90"] = rng.integers(0, 256, (90, 100), dtype=np.uint8)
    # Odd size with a bright diagonal structure and flat regions.
    img = np.full((77, 131), 90, np.float64)
    for k in range(0, 131, 9):
        img[:, k:k + 3] += 60
    img += _smooth_noise(rng, 77, 131, 5) * 0.3
    cases["stripes_131x77"] = np.clip(np.rint(img), 0, 255).astype(np.uint8)
    # Constant image: no corners at all.
    cases["constant_80x80"] = np.full((80, 80), 128, np.uint8)
    # Exactly at the detection size limit (2 * 31 + 1) and just below.
    cases["limit_63x63"] = rng.integers(0, 256, (63, 63), dtype=np.uint8)
    cases["small_62x70"] = rng.integers(0, 256, (70, 62), dtype=np.uint8)
    return cases


def s2s3_golden(ref_dir: Path):
    sys.path.insert(0, str(ref_dir))
    mod = importlib.import_module("tools.s2_s3.reference")
    out = {}
    for name, gray in gray_cases().items():
        for k in (0, 7, 50, 400):
            for thr in (10, 20):
                kp, desc = mod.detect_and_describe(gray, thr, k, 0.04)
                key = f"{name}__k{k}__t{thr}"
                out[name + "__gray"] = gray
                out[key + "__kp"] = kp.astype(np.float32)
                out[key + "__desc"] = desc.astype(np.uint8)
    # Blur and the building blocks that the tests also pin separately.
    g = gray_cases()["textured_160x120"]
    out["blur__textured_160x120"] = mod.blur(g)
    out["umax"] = mod.umax()
    out["taps"] = mod.BLUR_TAPS
    out["pattern"] = mod.load_pattern()
    return out


def _pattern(width, height, phase):
    yy, xx = np.indices((height, width), dtype=np.int32)
    return np.stack([(13 * xx + 7 * yy + phase) % 256,
                     (3 * xx + 17 * yy + 2 * phase) % 256,
                     ((9 * xx) ^ (5 * yy) ^ phase) % 256], axis=-1).astype(np.uint8)


def s6_golden(ref_dir: Path):
    sys.path.insert(0, str(ref_dir))
    mod = importlib.import_module("tools.s6_reference.reference")
    identity = np.eye(3)
    rng = np.random.default_rng(7)
    tex_a = np.clip(_smooth_noise(rng, 48, 64, 6)[..., None] + rng.uniform(-30, 30, (48, 64, 3)), 0, 255)
    tex_a = np.rint(tex_a).astype(np.uint8)
    tex_b = np.rint(np.clip(_smooth_noise(rng, 44, 60, 7)[..., None] + rng.uniform(-30, 30, (44, 60, 3)), 0, 255))
    tex_b = tex_b.astype(np.uint8)
    cases = [
        ("identity_odd", _pattern(13, 9, 11), _pattern(13, 9, 37), identity),
        ("negative_translation", _pattern(13, 9, 11), _pattern(11, 7, 37),
         np.array([[1, 0, -3], [0, 1, -2], [0, 0, 1.]])),
        ("fractional_translation", _pattern(13, 9, 11), _pattern(11, 7, 37),
         np.array([[1, 0, 0.375], [0, 1, -0.25], [0, 0, 1.]])),
        ("projective", _pattern(17, 13, 11), _pattern(15, 11, 37),
         np.array([[1.02, 0.03, -2], [-0.02, 0.98, 1], [0.001, -0.001, 1.]])),
        ("separated", _pattern(13, 9, 11), _pattern(11, 7, 37), np.array([[1, 0, 17], [0, 1, 2], [0, 0, 1.]])),
        ("single_pixel", _pattern(1, 1, 11), _pattern(1, 1, 37), identity),
        ("one_column", _pattern(1, 9, 11), _pattern(1, 9, 37), identity),
        ("one_row", _pattern(9, 1, 11), _pattern(9, 1, 37), identity),
        ("textured_overlap", tex_a, tex_b,
         np.array([[0.99, 0.05, 31.5], [-0.04, 1.01, 6.25], [0.0002, -0.0001, 1.]])),
    ]
    shallow = ("feather", "multiband_3", "multiband_5")
    cases = [(name, a, b, h, shallow) for name, a, b, h in cases]
    # Larger textured canvases (about 150 x 100, seven or more levels above
    # level 0) with a projective and a rotated map. The reference's NumPy
    # renderer implements feather, multiband_3 and multiband_5 only.

    def textured(h, w, cell):
        t = np.clip(_smooth_noise(rng, h, w, cell)[..., None] + rng.uniform(-30, 30, (h, w, 3)), 0, 255)
        return np.rint(t).astype(np.uint8)

    th = 0.3
    cases += [
        ("deep_projective", textured(80, 112, 9), textured(76, 108, 8),
         np.array([[1.01, 0.04, 47.25], [-0.03, 0.99, 9.5], [0.0003, -0.0002, 1.]]), shallow),
        ("deep_rotated", textured(96, 96, 10), textured(90, 100, 7),
         np.array([[np.cos(th), -np.sin(th), 58.0], [np.sin(th), np.cos(th), -21.5], [0.0, 0.0, 1.]]), shallow),
    ]
    out = {}
    for name, a, b, h, modes in cases:
        for mode in modes:
            r = mod.render(a, b, h, mode)
            g = r.geometry
            key = f"{name}__{mode}"
            out[key + "__a"] = a
            out[key + "__b"] = b
            out[key + "__h"] = np.asarray(h, np.float64)
            out[key + "__panorama"] = r.panorama
            out[key + "__mask"] = r.mask
            out[key + "__canvas"] = np.array([g.width, g.height, g.origin_x, g.origin_y, r.levels], np.int64)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--only", choices=("s2s3", "s6"), help="write one of the two files")
    args = ap.parse_args()
    ref = Path(args.reference)
    if args.only != "s6":
        np.savez_compressed(HERE / "s2s3_golden.npz", **s2s3_golden(ref))
        print("wrote", HERE / "s2s3_golden.npz")
    if args.only != "s2s3":
        np.savez_compressed(HERE / "s6_golden.npz", **s6_golden(ref))
        print("wrote", HERE / "s6_golden.npz")


if __name__ == "__main__":
    main()
