"""Oracle depth preprocessing: closed forms on planes, FP64 brute force on tiny images."""
import math

import numpy as np

from tests.helpers import intr, pixel_rays, plane_depth_mm, pre_params


def test_plane_maps_closed_form(orc):
    k = intr()
    D = 1.234
    m = orc.preprocess(k, pre_params(), plane_depth_mm(k, D), 3)
    X, Y = pixel_rays(k)
    lam = np.sqrt(X * X + Y * Y + 1.0)
    assert np.max(np.abs(m.raylen() - D * lam) / (D * lam)) < 5e-7
    for lvl in range(3):
        d = m.plane(lvl, "depth")
        assert np.max(np.abs(d - D)) <= 2 * np.spacing(np.float32(D))
        v = m.vertex(lvl)
        kl = orc.lib().kfo_intrinsics_level(k, lvl)
        Xl, Yl = pixel_rays(kl)
        ref = np.stack([Xl * D, Yl * D, np.full_like(Xl, D)], -1)
        assert np.max(np.abs(v - ref)) < 2e-6
        n = m.normal(lvl)
        inner = n[:-1, :-1]
        # 1-2 ulp depth differences over a 2.3 mm pixel baseline tilt normals by < 5e-4
        assert np.max(np.abs(inner - np.array([0.0, 0.0, -1.0]))) < 5e-4  # faces the camera
        assert np.all(np.isnan(n[-1, :, 0])) and np.all(np.isnan(n[:, -1, 0]))


def test_depth_range_gate_exact(orc):
    k = intr(64, 48)
    rng = np.random.default_rng(3)
    raw = rng.integers(0, 6000, size=(48, 64)).astype(np.uint16)
    p = pre_params(dmin=0.3, dmax=4.0)
    m = orc.preprocess(k, p, raw, 1)
    d = raw.astype(np.float32) * np.float32(0.001)
    valid = (d >= np.float32(0.3)) & (d <= np.float32(4.0))
    assert np.array_equal(m.raylen() > 0, valid)
    assert np.array_equal(~np.isnan(m.plane(0, "vz")), valid)
    assert np.array_equal(m.plane(0, "depth") > 0, valid)


def _bilateral_fp64(raw, k, R, ss, sr, dmin, dmax):
    H, W = raw.shape
    d = raw.astype(np.float64) * 0.001
    d[(d < dmin) | (d > dmax)] = 0.0
    out = np.zeros_like(d)
    for y in range(H):
        for x in range(W):
            c = d[y, x]
            if c == 0:
                continue
            sw = swd = 0.0
            for dy in range(-R, R + 1):
                for dx in range(-R, R + 1):
                    qy, qx = y + dy, x + dx
                    if 0 <= qy < H and 0 <= qx < W and d[qy, qx] > 0:
                        w = math.exp(-(dx * dx + dy * dy) / (2 * ss * ss) - (d[qy, qx] - c) ** 2 / (2 * sr * sr))
                        sw += w
                        swd += w * d[qy, qx]
            out[y, x] = swd / sw
    return out


def test_bilateral_matches_fp64_brute_force(orc):
    k = intr(24, 20)
    rng = np.random.default_rng(4)
    raw = (1500 + rng.normal(0, 25, size=(20, 24))).astype(np.uint16)
    raw[5:9, 10:14] = 2500  # a step edge
    raw[rng.random((20, 24)) < 0.1] = 0  # holes
    p = pre_params(radius=3, sigma_s=3.0, sigma_r=0.03)
    ref = _bilateral_fp64(raw, k, 3, 3.0, 0.03, 0.3, 4.0)
    m64 = orc.preprocess(k, p, raw, 1, precision=1)
    m32 = orc.preprocess(k, p, raw, 1, precision=0)
    assert np.max(np.abs(m64.plane(0, "depth") - ref)) < 1e-6  # (stored as float32)
    assert np.max(np.abs(m32.plane(0, "depth") - ref)) < 3e-6


def test_bilateral_invalid_pixels_never_contribute(orc):
    k = intr(32, 24)
    rng = np.random.default_rng(5)
    raw = (1200 + rng.normal(0, 20, size=(24, 32))).astype(np.uint16)
    raw[10, 10] = 0
    raw2 = raw.copy()
    raw2[10, 10] = 9000  # beyond max depth: still invalid
    a = orc.preprocess(k, pre_params(), raw, 1).plane(0, "depth")
    b = orc.preprocess(k, pre_params(), raw2, 1).plane(0, "depth")
    assert np.array_equal(a, b)


def test_bilateral_preserves_step_edge(orc):
    k = intr(40, 30)
    raw = np.full((30, 40), 1000, np.uint16)
    raw[:, 20:] = 2000  # 1 m step >> 3 sigma_r
    d = orc.preprocess(k, pre_params(sigma_r=0.03), raw, 1).plane(0, "depth")
    assert np.max(np.abs(d[:, :20] - 1.0)) < 1e-6
    assert np.max(np.abs(d[:, 20:] - 2.0)) < 1e-6


def test_pyramid_gated_block_average(orc):
    k = intr(32, 24)
    rng = np.random.default_rng(6)
    raw = (1500 + rng.normal(0, 30, size=(24, 32))).astype(np.uint16)
    raw[rng.random((24, 32)) < 0.15] = 0
    raw[0:6, 0:6] = 3000
    p = pre_params(radius=2, gate=0.05)
    m = orc.preprocess(k, p, raw, 3)
    for lvl in (1, 2):
        fine = m.plane(lvl - 1, "depth").astype(np.float64)
        co = m.plane(lvl, "depth")
        H, W = co.shape
        for y in range(H):
            for x in range(W):
                c = fine[2 * y, 2 * x]
                if c == 0:
                    assert co[y, x] == 0
                    continue
                blk = fine[2 * y:2 * y + 2, 2 * x:2 * x + 2].reshape(-1)
                sel = blk[(blk > 0) & (np.abs(blk - c) < 0.05)]
                assert abs(co[y, x] - sel.mean()) < 2e-6


def test_normals_of_tilted_plane(orc):
    k = intr(80, 60, fx=70.0, fy=70.0)
    # plane n.p = c with n = normalized(0.3, -0.2, -1) facing the camera
    nrm = np.array([0.3, -0.2, -1.0])
    nrm /= np.linalg.norm(nrm)
    c = -1.5
    X, Y = pixel_rays(k)
    z = c / (nrm[0] * X + nrm[1] * Y + nrm[2])
    raw = np.round(z * 1000).astype(np.uint16)
    m = orc.preprocess(k, pre_params(radius=0, sigma_s=1.0), raw, 1)
    n = m.normal(0)[:-1, :-1].reshape(-1, 3)
    cosang = n @ nrm
    assert np.median(cosang) > 0.999  # mm quantisation perturbs individual pixels
