"""Pins the prior's FP64 oracle: shapes and parameter count, im2col/col2im
adjointness, and its backward against central finite differences."""
import numpy as np
import pytest

from oracle import prior_oracle as po


def _batch(rng, B=2, F=3):
    frames = rng.uniform(0.5, 3.0, size=(F, po.H0, po.W0))
    frames[rng.random(frames.shape) < 0.1] = 0.0
    pairs = rng.integers(0, F, size=(B, 2))
    prev = rng.normal(0, 0.5, size=(B, 6))
    target = rng.normal(0, 0.5, size=(B, 6))
    return frames, pairs, prev, target


def test_layout_and_shapes(kf):
    hs, ws = po.shapes()
    assert hs == [120, 60, 30, 15, 8] and ws == [160, 80, 40, 20, 10]
    _, total = po.layout()
    assert total == 165990
    assert kf.lib().kf_prior_param_count() == total


def test_im2col_col2im_adjoint():
    rng = np.random.default_rng(0)
    X = rng.normal(size=(2, 15, 20, 3))
    col = po.im2col(X)
    D = rng.normal(size=col.shape)
    # <im2col(X), D> == <X, col2im(D)>
    assert abs((col * D).sum() - (X * po.col2im(D, 2, 15, 20, 3)).sum()) < 1e-9


def test_im2col_matches_direct_convolution():
    rng = np.random.default_rng(1)
    X = rng.normal(size=(1, 9, 11, 2))
    Wt = rng.normal(size=(4, 18))
    Y = (po.im2col(X) @ Wt.T).reshape(1, 5, 6, 4)
    Wk = Wt.reshape(4, 3, 3, 2)
    for oy in range(5):
        for ox in range(6):
            acc = np.zeros(4)
            for ky in range(3):
                for kx in range(3):
                    iy, ix = 2 * oy + ky - 1, 2 * ox + kx - 1
                    if 0 <= iy < 9 and 0 <= ix < 11:
                        acc += Wk[:, ky, kx, :] @ X[0, iy, ix]
            assert np.allclose(Y[0, oy, ox], acc)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(2)
    frames, pairs, prev, target = _batch(rng)
    p = po.init_params(3)
    loss, g = po.gradients(p, frames, pairs, prev, target, bf=False)
    L, total = po.layout()
    idx = []
    for wo, ws, bo, bl in L:
        idx += list(rng.integers(wo, wo + ws[0] * ws[1], 4)) + [bo + int(rng.integers(0, bl))]
    h = 1e-6
    checked = 0
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = po.gradients(pp, frames, pairs, prev, target, bf=False)
        lm, _ = po.gradients(pm, frames, pairs, prev, target, bf=False)
        fd = (lp - lm) / (2 * h)
        if abs(g[i]) < 1e-9 and abs(fd) < 1e-9:
            continue
        assert abs(fd - g[i]) <= 1e-5 * max(abs(fd), abs(g[i])) + 1e-9, (i, fd, g[i])
        checked += 1
    assert checked >= 20


def test_bf16_rounding_is_rne():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.1415926], dtype=np.float32)
    got = po.bf16(x)
    assert got[0] == 1.0
    assert got[1] == 1.0  # halfway between 1 and 1 + 2^-7: ties to the even mantissa
    assert got[2] == 1.0 + 2 ** -6  # tie -> even (...1.0000010)
    assert got[3] == 1.0
    assert got[4] == -2.5
    assert abs(got[5] - 3.140625) < 1e-12
