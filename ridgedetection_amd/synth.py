"""Seeded synthetic point clouds for ridge detection.

The reference generates its test clouds with `std::random_device` seeds
(ridge_detection/rd/cpu/samples_generator.hpp:42-170), so its inputs are not
reproducible.  These generators draw the same shapes from numpy's PCG64 with an
explicit seed.  All functions return ROW_MAJOR arrays of shape (n, dim).
"""

from __future__ import annotations

import math

import numpy as np


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def spiral2d(n: int, seed: int = 0, sigma: float = 0.1, turns: float = 3.0,
             dtype=np.float32) -> np.ndarray:
    """Archimedean spiral: t ~ U[0, 2*pi*turns], (0.5 t cos t, 0.5 t sin t) + N(0, sigma^2)."""
    g = _rng(seed)
    t = g.uniform(0.0, 2.0 * math.pi * turns, n)
    x = 0.5 * t * np.cos(t)
    y = 0.5 * t * np.sin(t)
    p = np.stack([x, y], axis=1) + g.normal(0.0, sigma, (n, 2))
    return np.ascontiguousarray(p.astype(dtype))


def helix3d(n: int, seed: int = 0, sigma: float = 0.05, turns: float = 5.0,
            pitch: float = 0.15, dtype=np.float32) -> np.ndarray:
    """Helix: t ~ U[0, 2*pi*turns], (cos t, sin t, pitch t) + N(0, sigma^2)."""
    g = _rng(seed)
    t = g.uniform(0.0, 2.0 * math.pi * turns, n)
    p = np.stack([np.cos(t), np.sin(t), pitch * t], axis=1) + g.normal(0.0, sigma, (n, 3))
    return np.ascontiguousarray(p.astype(dtype))


def helix_curve(t: np.ndarray, pitch: float = 0.15) -> np.ndarray:
    """Noise-free helix points at parameters t."""
    return np.stack([np.cos(t), np.sin(t), pitch * t], axis=-1)


def helix_loss(S: np.ndarray, turns: float = 5.0, pitch: float = 0.15):
    """Sum of squared distances of the points S (K x 3) to the noise-free helix
    t in [0, 2 pi turns], and its gradient with respect to S (both float64).

    The nearest parameter of every point comes from a dense grid refined by Newton steps on
    d/dt |s - h(t)|^2 / 2 = x sin t - y cos t + pitch (pitch t - z); at that parameter the
    gradient of the squared distance is 2 (s - h(t*)) (the t-derivative vanishes there, or
    t* sits on an end of the parameter range)."""
    S = np.asarray(S, dtype=np.float64)
    t_hi = 2.0 * math.pi * turns
    grid = np.linspace(0.0, t_hi, int(t_hi / 0.05) + 1)
    H = helix_curve(grid, pitch)
    t = np.empty(len(S))
    for a in range(0, len(S), 256):
        blk = S[a:a + 256]
        d2 = ((blk[:, None, :] - H[None, :, :]) ** 2).sum(-1)
        t[a:a + 256] = grid[np.argmin(d2, axis=1)]
    x, y, z = S[:, 0], S[:, 1], S[:, 2]
    for _ in range(12):
        g = x * np.sin(t) - y * np.cos(t) + pitch * (pitch * t - z)
        h = x * np.cos(t) + y * np.sin(t) + pitch * pitch
        step = np.where(h > 1e-12, g / np.where(h > 1e-12, h, 1.0), 0.0)
        t = np.clip(t - step, 0.0, t_hi)
    diff = S - helix_curve(t, pitch)
    return float((diff * diff).sum()), 2.0 * diff


def bezier_union(n: int, dim: int = 6, curves: int = 4, seed: int = 0, sigma: float = 0.02,
                 outlier_frac: float = 0.05, outlier_box: float = 1.2,
                 dtype=np.float32) -> np.ndarray:
    """Union of cubic Bezier curves with control points ~U[-1,1]^dim, isotropic noise and
    a fraction of uniform outliers in [-outlier_box, outlier_box]^dim.  Points are shuffled."""
    g = _rng(seed)
    ctrl = g.uniform(-1.0, 1.0, (curves, 4, dim))
    n_out = int(round(n * outlier_frac))
    n_in = n - n_out
    which = g.integers(0, curves, n_in)
    u = g.uniform(0.0, 1.0, n_in)[:, None]
    c = ctrl[which]
    b = ((1 - u) ** 3) * c[:, 0] + 3 * ((1 - u) ** 2) * u * c[:, 1] \
        + 3 * (1 - u) * (u ** 2) * c[:, 2] + (u ** 3) * c[:, 3]
    b = b + g.normal(0.0, sigma, (n_in, dim))
    o = g.uniform(-outlier_box, outlier_box, (n_out, dim))
    p = np.concatenate([b, o], axis=0)
    g.shuffle(p, axis=0)
    return np.ascontiguousarray(p.astype(dtype))


def bezier_union_curves(dim: int = 6, curves: int = 4, seed: int = 0, per_curve: int = 50000) -> np.ndarray:
    """Noise-free points of the curves bezier_union(..., dim, curves, seed) samples around (the same
    control points: they are the generator's first draw), per_curve evenly spaced parameters each."""
    g = _rng(seed)
    ctrl = g.uniform(-1.0, 1.0, (curves, 4, dim))
    u = np.linspace(0.0, 1.0, per_curve)[:, None]
    pts = [((1 - u) ** 3) * c[0] + 3 * ((1 - u) ** 2) * u * c[1] + 3 * (1 - u) * (u ** 2) * c[2] + (u ** 3) * c[3]
           for c in ctrl]
    return np.ascontiguousarray(np.concatenate(pts, axis=0))


def segment_chain(n: int, dim: int, seed: int = 0, segments: int = 4, sigma: float = 0.05,
                  dtype=np.float32) -> np.ndarray:
    """Polyline through `segments+1` random vertices in [-1,1]^dim with noise."""
    g = _rng(seed)
    v = g.uniform(-1.0, 1.0, (segments + 1, dim))
    which = g.integers(0, segments, n)
    u = g.uniform(0.0, 1.0, n)[:, None]
    p = v[which] * (1 - u) + v[which + 1] * u + g.normal(0.0, sigma, (n, dim))
    return np.ascontiguousarray(p.astype(dtype))


def batched_clouds(count: int = 512, seed: int = 0, min_pts: int = 2000, max_pts: int = 500000,
                   dtype=np.float32):
    """Clouds of uneven sizes: half with d=2, half with d=3; sizes log-uniform in
    [min_pts, max_pts]; each a spiral, helix or segment chain with noise
    sigma ~ U[0.01, 0.1] * extent and radius r = 4 sigma.  Returns a list of
    (points, r1, r2) with r2 = 2 r1."""
    g = _rng(seed)
    out = []
    for i in range(count):
        dim = 2 if i < count // 2 else 3
        n = int(round(math.exp(g.uniform(math.log(min_pts), math.log(max_pts)))))
        kind = int(g.integers(0, 3))
        frac = float(g.uniform(0.01, 0.1))
        s = int(g.integers(0, 2 ** 31))
        if kind == 0:
            base = spiral2d(n, s, sigma=0.0, dtype=np.float64) if dim == 2 else \
                helix3d(n, s, sigma=0.0, dtype=np.float64)
        elif kind == 1:
            base = helix3d(n, s, sigma=0.0, dtype=np.float64)[:, :dim]
        else:
            base = segment_chain(n, dim, s, sigma=0.0, dtype=np.float64)
        extent = float(np.max(base.max(axis=0) - base.min(axis=0)))
        sigma = frac * extent
        pts = base + _rng(s + 1).normal(0.0, sigma, base.shape)
        r1 = 4.0 * sigma
        out.append((np.ascontiguousarray(pts.astype(dtype)), r1, 2.0 * r1))
    return out


def hausdorff(a: np.ndarray, b: np.ndarray) -> float:
    """Symmetric Hausdorff distance between two point sets (rows), as
    ridge_detection/rd/utils/assessment_quality.hpp:91-220 defines it."""
    if len(a) == 0 or len(b) == 0:
        return -1.0
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)

    def one_way(x, y):
        best = np.full(len(x), np.inf)
        for s in range(0, len(y), 4096):
            d = ((x[:, None, :] - y[None, s:s + 4096, :]) ** 2).sum(-1)
            best = np.minimum(best, d.min(axis=1))
        return float(np.sqrt(best.max()))

    return max(one_way(a, b), one_way(b, a))
