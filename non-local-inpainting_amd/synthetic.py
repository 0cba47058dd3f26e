"""Seeded synthetic inputs for the inpainting workloads.

There are no reference datasets with expected outputs, so every workload is
generated here from a fixed seed:

* texture: per-channel linear ramp ``a_c + b_c*x/W + c_c*y/H`` plus a sum of K
  sinusoids of amplitude 0.2 (random orientation, wavelength and phase, a
  per-channel gain) plus Gaussian noise sigma 0.02, clamped to [0, 1];
* depth: random planar segments in 0.5-10 m plus Gaussian noise sigma 5 mm;
* holes: a centred square, a union of random ellipses scaled to a target area
  fraction, or random vertical/diagonal stripes.

Images are float32 HWC (index ``C*(W*y+x)+c``), masks uint8 0/1 (1 = hole).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def texture(width: int, height: int, channels: int = 3, sinusoids: int = 3,
            seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:height, 0:width].astype(np.float64)
    out = np.empty((height, width, channels), dtype=np.float64)
    a = rng.uniform(0.25, 0.55, channels)
    b = rng.uniform(-0.2, 0.2, channels)
    c = rng.uniform(-0.2, 0.2, channels)
    for ch in range(channels):
        out[..., ch] = a[ch] + b[ch] * x / width + c[ch] * y / height
    for _ in range(sinusoids):
        theta = rng.uniform(0.0, math.pi)
        lam = rng.uniform(8.0, 96.0)
        phi = rng.uniform(0.0, 2 * math.pi)
        gain = rng.uniform(0.5, 1.0, channels)
        wave = 0.2 * np.sin(2 * math.pi * (x * math.cos(theta) + y * math.sin(theta)) / lam + phi)
        for ch in range(channels):
            out[..., ch] += gain[ch] * wave
    out += rng.normal(0.0, 0.02, out.shape)
    return np.clip(out, 0.0, 1.0).astype(np.float32)


def depth_planes(width: int, height: int, segments: int = 5, seed: int = 0) -> np.ndarray:
    """Depth in metres: a Voronoi partition into planar segments in [0.5, 10]."""
    rng = np.random.default_rng(seed + 7)
    y, x = np.mgrid[0:height, 0:width].astype(np.float64)
    cx = rng.uniform(0, width, segments)
    cy = rng.uniform(0, height, segments)
    d2 = (x[None] - cx[:, None, None]) ** 2 + (y[None] - cy[:, None, None]) ** 2
    label = np.argmin(d2, axis=0)
    out = np.zeros((height, width), dtype=np.float64)
    for s in range(segments):
        z0 = rng.uniform(1.0, 9.0)
        gx = rng.uniform(-1.0, 1.0) / width
        gy = rng.uniform(-1.0, 1.0) / height
        plane = z0 + gx * (x - cx[s]) * 4 + gy * (y - cy[s]) * 4
        out[label == s] = plane[label == s]
    out = np.clip(out, 0.5, 10.0) + rng.normal(0.0, 0.005, out.shape)
    return out.astype(np.float32)


def square_hole(width: int, height: int, side: int) -> np.ndarray:
    m = np.zeros((height, width), dtype=np.uint8)
    x0 = (width - side) // 2
    y0 = (height - side) // 2
    m[y0:y0 + side, x0:x0 + side] = 1
    return m


def ellipse_holes(width: int, height: int, count: int, fraction: float, seed: int = 0,
                  border: int = 8) -> np.ndarray:
    """Union of ``count`` random rotated ellipses, scaled until the hole covers
    ``fraction`` of the image (within 0.5 %). Ellipses stay ``border`` px inside."""
    rng = np.random.default_rng(seed + 13)
    y, x = np.mgrid[0:height, 0:width].astype(np.float64)
    params = []
    for _ in range(count):
        params.append((rng.uniform(0.2, 0.8) * width, rng.uniform(0.2, 0.8) * height,
                       rng.uniform(0.5, 1.5), rng.uniform(0.5, 1.5), rng.uniform(0, math.pi)))

    def build(scale: float) -> np.ndarray:
        m = np.zeros((height, width), dtype=bool)
        base = math.sqrt(width * height) * scale
        for cx, cy, ax, ay, th in params:
            dx = x - cx
            dy = y - cy
            u = (dx * math.cos(th) + dy * math.sin(th)) / (base * ax)
            v = (-dx * math.sin(th) + dy * math.cos(th)) / (base * ay)
            m |= (u * u + v * v) <= 1.0
        m[:border, :] = False
        m[-border:, :] = False
        m[:, :border] = False
        m[:, -border:] = False
        return m

    lo, hi = 0.001, 1.0
    m = build(hi)
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        m = build(mid)
        f = m.mean()
        if abs(f - fraction) < 0.005:
            break
        if f < fraction:
            lo = mid
        else:
            hi = mid
    return m.astype(np.uint8)


def stripe_holes(width: int, height: int, count: int = 3, stripe_width: int = 40,
                 seed: int = 0) -> np.ndarray:
    """Random vertical or diagonal stripes crossing the whole image (they touch
    the image border)."""
    rng = np.random.default_rng(seed + 29)
    y, x = np.mgrid[0:height, 0:width].astype(np.float64)
    m = np.zeros((height, width), dtype=bool)
    for k in range(count):
        x0 = rng.uniform(0.3, 0.7) * width
        # the first stripe is vertical, the others diagonal (1.6-2.4 px across per px
        # down), long enough that three 40 px stripes cover ~12 % of a 1920x1080 image
        slope = 0.0 if k == 0 else rng.choice([-1.0, 1.0]) * rng.uniform(1.6, 2.4) * height / width * 1920 / 1080
        xc = x0 + slope * (y - height / 2)
        m |= np.abs(x - xc) * math.cos(math.atan(slope)) <= stripe_width / 2
    return m.astype(np.uint8)


@dataclass
class Workload:
    name: str
    image: np.ndarray           # (H, W, C) float32
    mask: np.ndarray            # (H, W) uint8
    truth: np.ndarray           # ground truth image before the hole is erased
    params: dict = field(default_factory=dict)


def config_small(seed: int = 0) -> Workload:
    """256x256 RGB, centred 64x64 hole, nlmeans, patch 9, psigma 1e4, 4 scales,
    coarse 0.1, 10 iterations, Poisson initialisation."""
    img = texture(256, 256, 3, 3, seed)
    mask = square_hole(256, 256, 64)
    return Workload("rgb256_square64_nlmeans", img, mask, img.copy(), dict(
        method="nlmeans", patch=9, sigma=10000.0, scales=4, coarse=0.1, iters=10, init="poisson"))


def config_1024(seed: int = 1) -> Workload:
    """1024x1024 RGB, union of 6 ellipses ~15 %, nlpoisson lambda 0.05, patch 9,
    11 scales, coarse 0.1, <=200 iterations, CG tol 1e-4 <=100 iterations."""
    img = texture(1024, 1024, 3, 8, seed)
    mask = ellipse_holes(1024, 1024, 6, 0.15, seed)
    return Workload("rgb1024_ellipses15_nlpoisson", img, mask, img.copy(), dict(
        method="nlpoisson", patch=9, sigma=10000.0, lam=0.05, scales=11, coarse=0.1, iters=200,
        cg_tol=1e-4, cg_limit=100, init="poisson"))


def config_rgbd(seed: int = 2) -> Workload:
    """1920x1080 RGB + depth, 3 stripes 40 px wide, nldepth lambdar 1 lambdad 0,
    patch 9, 11 scales."""
    rgb = texture(1920, 1080, 3, 8, seed)
    dep = depth_planes(1920, 1080, 5, seed)
    img = np.concatenate([rgb, dep[..., None]], axis=2)
    mask = stripe_holes(1920, 1080, 3, 40, seed)
    return Workload("rgbd1920x1080_stripes_nldepth", img, mask, img.copy(), dict(
        method="nldepth", patch=9, sigma=10000.0, lambda_rgb=1.0, lambda_d=0.0, scales=11,
        coarse=0.1, iters=200, init="poisson"))


def scaled(w: Workload, width: int, height: int, seed: int = 0) -> Workload:
    """A smaller instance of the same workload family (for parity tests)."""
    c = w.image.shape[2]
    img = texture(width, height, 3, 3, seed)
    if c == 4:
        img = np.concatenate([img, depth_planes(width, height, 5, seed)[..., None]], axis=2)
    if "square" in w.name:
        mask = square_hole(width, height, max(4, width // 4))
    elif "stripes" in w.name:
        mask = stripe_holes(width, height, 3, max(4, width // 48), seed)
    else:
        mask = ellipse_holes(width, height, 6, 0.15, seed)
    return Workload(w.name + f"_{width}x{height}", img, mask, img.copy(), dict(w.params))


def config_uneven(count: int = 64, seed: int = 3, min_side: int = 128, max_side: int = 1024) -> list:
    """Uneven batch: ``count`` RGB images with sides drawn uniformly from
    [min_side, max_side] (independent width and height), random-ellipse holes of
    5-30 % area, alternately nlmeans and nlmedians, patch 7 or 9, 8 scales.  The
    coarsest size ratio is 0.1 unless that would leave the coarsest level shorter
    than 48 px, where no complete source patch survives around a hole (the
    reference stops with "Empty source mask" there); then it is 48 / min(W, H).
    Returns a list of Workload."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        W = int(rng.integers(min_side, max_side + 1))
        H = int(rng.integers(min_side, max_side + 1))
        frac = float(rng.uniform(0.05, 0.30))
        img = texture(W, H, 3, 5, seed * 1000 + k)
        mask = ellipse_holes(W, H, int(rng.integers(2, 6)), frac, seed * 1000 + k)
        method = "nlmeans" if k % 2 == 0 else "nlmedians"
        patch = 7 if rng.uniform() < 0.5 else 9
        coarse = max(0.1, 48.0 / min(W, H))
        out.append(Workload(f"uneven{k}_{W}x{H}_{method}_p{patch}", img, mask, img.copy(), dict(
            method=method, patch=patch, sigma=10000.0, scales=8, coarse=coarse, iters=200, init="poisson",
            seed=k + 1)))
    return out
