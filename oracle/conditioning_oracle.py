"""Numpy restatement of the camera-side depth conditioning.  TEST INFRASTRUCTURE ONLY.

median5: 5x5 median of the non-null neighbours (pixels outside the image skipped),
null when at most 12 are valid, integer mean of the middle pair for an even count.
undistort: Brown-Conrady forward model from the undistorted normalized coordinates,
bilinear sampling requiring all four source pixels valid, rounding to nearest,
evaluated in float32 in the device's operation order.
"""
from __future__ import annotations

import numpy as np


def median5(d: np.ndarray, null: int = 0) -> np.ndarray:
    H, W = d.shape
    pad = np.full((H + 4, W + 4), -1, np.int64)
    pad[2:H + 2, 2:W + 2] = d
    win = np.stack([pad[2 + j:2 + j + H, 2 + i:2 + i + W] for j in range(-2, 3) for i in range(-2, 3)], -1)
    valid = (win >= 0) & (win != null)
    n = valid.sum(-1)
    big = np.where(valid, win, 1 << 20)
    s = np.sort(big, -1)
    idx = n // 2
    hi = np.take_along_axis(s, idx[..., None], -1)[..., 0]
    lo = np.take_along_axis(s, np.maximum(idx - 1, 0)[..., None], -1)[..., 0]
    med = np.where(n % 2 == 1, hi, (hi + lo) // 2)
    return np.where(n > 12, med, null).astype(np.uint16)


def undistort(d: np.ndarray, fx, fy, cx, cy, k1, k2, p1, p2) -> np.ndarray:
    f = np.float32
    H, W = d.shape
    x, y = np.meshgrid(np.arange(W, dtype=f), np.arange(H, dtype=f))
    fx, fy, cx, cy, k1, k2, p1, p2 = (f(v) for v in (fx, fy, cx, cy, k1, k2, p1, p2))
    xu = (x - cx) / fx
    yu = (y - cy) / fy
    r2 = xu * xu + yu * yu
    radial = f(1.0) + (k1 * r2 + (k2 * r2) * r2)
    xd = xu * radial + ((f(2.0) * p1) * xu * yu + p2 * (r2 + (f(2.0) * xu) * xu))
    yd = yu * radial + ((f(2.0) * p2) * xu * yu + p1 * (r2 + (f(2.0) * yu) * yu))
    u = fx * xd + cx
    v = fy * yd + cy
    fu, fv = np.floor(u), np.floor(v)
    ok = (fu >= 0) & (fv >= 0) & (fu <= W - 2) & (fv <= H - 2)
    iu = np.where(ok, fu, 0).astype(np.int64)
    iv = np.where(ok, fv, 0).astype(np.int64)
    a, b = d[iv, iu].astype(f), d[iv, iu + 1].astype(f)
    c, e = d[iv + 1, iu].astype(f), d[iv + 1, iu + 1].astype(f)
    ok &= (a > 0) & (b > 0) & (c > 0) & (e > 0)
    tu, tv = u - fu, v - fv
    top = a + (b - a) * tu
    bot = c + (e - c) * tu
    r = (top + (bot - top) * tv + f(0.5)).astype(np.uint16)
    return np.where(ok, r, 0).astype(np.uint16)
