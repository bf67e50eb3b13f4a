"""Pins the conditioning restatement: brute-force median, identity undistortion."""
import numpy as np

from oracle import conditioning_oracle as co


def test_median_vs_brute_force():
    rng = np.random.default_rng(1)
    d = rng.integers(0, 3000, size=(17, 23)).astype(np.uint16)
    d[rng.random(d.shape) < 0.35] = 0
    got = co.median5(d)
    H, W = d.shape
    for y in range(H):
        for x in range(W):
            vals = sorted(int(d[yy, xx]) for yy in range(max(0, y - 2), min(H, y + 3))
                          for xx in range(max(0, x - 2), min(W, x + 3)) if d[yy, xx] != 0)
            n = len(vals)
            exp = 0 if n <= 12 else (vals[n // 2] if n % 2 else (vals[n // 2] + vals[n // 2 - 1]) // 2)
            assert got[y, x] == exp


def test_undistort_identity_and_shift():
    d = (np.arange(40 * 30) % 900 + 500).astype(np.uint16).reshape(30, 40)
    out = co.undistort(d, 30.0, 30.0, 19.5, 14.5, 0, 0, 0, 0)
    assert np.array_equal(out[:-1, :-1], d[:-1, :-1])
    # against a scalar FP64 evaluation of the Brown-Conrady model and bilinear sampling
    rng = np.random.default_rng(3)
    d3 = rng.integers(800, 2500, size=(30, 40)).astype(np.uint16)
    K = (30.0, 31.0, 19.5, 14.5, 0.08, -0.03, 0.002, -0.001)
    out3 = co.undistort(d3, *K)
    fx, fy, cx, cy, k1, k2, p1, p2 = K
    for y in range(30):
        for x in range(40):
            xu, yu = (x - cx) / fx, (y - cy) / fy
            r2 = xu * xu + yu * yu
            rad = 1 + k1 * r2 + k2 * r2 * r2
            u = fx * (xu * rad + 2 * p1 * xu * yu + p2 * (r2 + 2 * xu * xu)) + cx
            v = fy * (yu * rad + 2 * p2 * xu * yu + p1 * (r2 + 2 * yu * yu)) + cy
            iu, iv = int(np.floor(u)), int(np.floor(v))
            if not (0 <= iu <= 38 and 0 <= iv <= 28):
                assert out3[y, x] == 0
                continue
            tu, tv = u - iu, v - iv
            a, b, c, e = (float(d3[iv, iu]), float(d3[iv, iu + 1]), float(d3[iv + 1, iu]), float(d3[iv + 1, iu + 1]))
            ref = (a + (b - a) * tu) * (1 - tv) + (c + (e - c) * tu) * tv
            assert abs(int(out3[y, x]) - ref) <= 0.51 + 1e-3
