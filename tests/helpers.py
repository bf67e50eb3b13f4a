"""Shared test helpers: analytic inputs and FP64 brute-force references."""
from __future__ import annotations

import math

import numpy as np

from openfusion_amd import kinfu as kf

K640 = dict(fx=525.0, fy=525.0, cx=319.5, cy=239.5, width=640, height=480)


def intr(width=640, height=480, fx=525.0, fy=525.0, cx=None, cy=None) -> kf.Intrinsics:
    cx = (width - 1) / 2 if cx is None else cx
    cy = (height - 1) / 2 if cy is None else cy
    return kf.Intrinsics(fx, fy, cx, cy, width, height)


def pre_params(radius=3, sigma_s=3.0, sigma_r=0.03, gate=0.09, dmin=0.3, dmax=4.0) -> kf.PreprocessParams:
    return kf.PreprocessParams(0.001, dmin, dmax, radius, sigma_s, sigma_r, gate)


def integ_params(dmin=0.3, dmax=4.0) -> kf.IntegrateParams:
    return kf.IntegrateParams(dmin, dmax)


def ray_params(dmin=0.3, far=0.8, near=0.5, max_steps=2048) -> kf.RaycastParams:
    return kf.RaycastParams(dmin, far, near, max_steps)


def icp_params(levels=3, iters=(10, 5, 4), dist=0.1, deg=20.0, comp=1, single=0) -> kf.IcpParams:
    p = kf.IcpParams()
    p.levels = levels
    for i in range(3):
        p.iterations[i] = iters[i] if i < len(iters) else 0
    p.dist_thresh_m, p.normal_thresh_deg, p.compensated = dist, deg, comp
    p.single_launch = single
    return p


def vol_desc(n=64, extent=2.0, trunc=0.05, maxw=128.0, fmt=kf.KF_VOXEL_F32, origin=None) -> kf.VolumeDesc:
    d = kf.VolumeDesc()
    d.n, d.extent_m, d.trunc_m, d.max_weight, d.format = n, extent, trunc, maxw, fmt
    o = origin if origin is not None else (-extent / 2,) * 3
    for i in range(3):
        d.origin_m[i] = o[i]
    return d


def identity() -> np.ndarray:
    return np.eye(4)


def pose(rx=0.0, ry=0.0, rz=0.0, t=(0.0, 0.0, 0.0)) -> np.ndarray:
    """T_wc = [Rz Ry Rx | t] (radians)."""
    cx, sx, cy, sy, cz, sz = math.cos(rx), math.sin(rx), math.cos(ry), math.sin(ry), math.cos(rz), math.sin(rz)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    T = np.eye(4)
    T[:3, :3] = Rz @ Ry @ Rx
    T[:3, 3] = t
    return T


def plane_depth_mm(k: kf.Intrinsics, D: float) -> np.ndarray:
    return np.full((k.height, k.width), int(round(D * 1000)), dtype=np.uint16)


def pixel_rays(k: kf.Intrinsics) -> tuple[np.ndarray, np.ndarray]:
    x = np.arange(k.width, dtype=np.float64)
    y = np.arange(k.height, dtype=np.float64)
    xn = (x - k.cx) / k.fx
    yn = (y - k.cy) / k.fy
    return np.meshgrid(xn, yn)


def voxel_centres(d: kf.VolumeDesc) -> np.ndarray:
    n = d.n
    h = d.extent_m / n
    c = (np.arange(n) + 0.5) * h
    X, Y, Z = np.meshgrid(c + d.origin_m[0], c + d.origin_m[1], c + d.origin_m[2], indexing="ij")
    return np.stack([X, Y, Z], -1)


def fill_volume(d: kf.VolumeDesc, f: np.ndarray, w: np.ndarray | float = 1.0) -> np.ndarray:
    """Raw voxel bytes (kf layout) from tsdf/T and weights (n,n,n)."""
    f = np.asarray(f, np.float32)
    w = np.broadcast_to(np.asarray(w, np.float32), f.shape).reshape(-1)
    f = f.reshape(-1)
    if d.format == kf.KF_VOXEL_F16:
        lo = f.astype(np.float16).view(np.uint16).astype(np.uint32)
        hi = w.astype(np.uint32) << 16
        return (lo | hi).view(np.uint8)
    fw = np.empty((f.size, 2), np.float32)
    fw[:, 0], fw[:, 1] = f, w
    return fw.reshape(-1).view(np.uint8)


def f64_volume(f: np.ndarray, w: np.ndarray | float = 1.0) -> np.ndarray:
    f = np.asarray(f, np.float64)
    w = np.broadcast_to(np.asarray(w, np.float64), f.shape).reshape(-1)
    f = f.reshape(-1)
    fw = np.empty((f.size, 2), np.float64)
    fw[:, 0], fw[:, 1] = f, w
    return fw.reshape(-1).view(np.uint8)


def decode_volume(d: kf.VolumeDesc, raw: np.ndarray, fmt: int | None = None):
    fmt = d.format if fmt is None else fmt
    n = d.n
    if fmt == kf.KF_VOXEL_F16:
        w32 = raw.view(np.uint32)
        f = (w32 & 0xFFFF).astype(np.uint16).view(np.float16).astype(np.float32)
        w = (w32 >> 16).astype(np.float32)
    elif fmt == kf.KF_VOXEL_F32:
        fw = raw.view(np.float32).reshape(-1, 2)
        f, w = fw[:, 0], fw[:, 1]
    else:
        fw = raw.view(np.float64).reshape(-1, 2)
        f, w = fw[:, 0], fw[:, 1]
    return f.reshape(n, n, n), w.reshape(n, n, n)


def mask_bits(mask_words: np.ndarray, count: int) -> np.ndarray:
    b = np.unpackbits(mask_words.view(np.uint8), bitorder="little")
    return b[:count].astype(bool)


def same_bits(a: np.ndarray, b: np.ndarray) -> bool:
    """Bitwise equality of float arrays treating all NaNs as equal."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False
    return np.array_equal(a[~na].view(np.uint32), b[~nb].view(np.uint32))
