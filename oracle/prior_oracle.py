"""FP64 numpy oracle of the learned pose prior.  TEST INFRASTRUCTURE ONLY.

Restates, independently of the HIP kernels, the network the product trains and
runs (two 160x120 depth maps -> four 3x3/s2/p1 convolutions 2->32->64->64->128 with
ReLU -> global average pool -> concat(previous twist) -> FC 134->256 ReLU -> FC
256->6, plus the previous twist as a skip connection) and the gradient of
mean((out - target)^2).  With ``bf16=True`` every GEMM
operand is rounded to bfloat16 (round to nearest even) at the same points as the
device (im2col columns, weights, FC inputs, output gradients), with products
accumulated in FP64, so the device's FP32-accumulated results differ only by
accumulation rounding.  With ``bf16=False`` it is the plain FP64 network used for
finite-difference checks.

Parameter blob: per layer W (rows = output channels; conv columns ordered
(ky, kx, cin)) then b, for conv1..conv4, fc1, fc2.
"""
from __future__ import annotations

import numpy as np

H0, W0 = 120, 160
CONVS = [(2, 32), (32, 64), (64, 64), (64, 128)]
FEAT, TWIST, HIDDEN = 128, 6, 256
FC1_IN = FEAT + TWIST


def conv_out(n: int) -> int:
    return (n + 2 - 3) // 2 + 1


def shapes():
    hs, ws = [H0], [W0]
    for _ in CONVS:
        hs.append(conv_out(hs[-1]))
        ws.append(conv_out(ws[-1]))
    return hs, ws


def layout():
    """[(w_off, w_shape, b_off, b_len)] for the 6 layers, and the total count."""
    out, o = [], 0
    for cin, cout in CONVS:
        ws = (cout, 9 * cin)
        out.append((o, ws, o + ws[0] * ws[1], cout))
        o += ws[0] * ws[1] + cout
    for nin, nout in ((FC1_IN, HIDDEN), (HIDDEN, TWIST)):
        out.append((o, (nout, nin), o + nout * nin, nout))
        o += nout * nin + nout
    return out, o


def unpack(params: np.ndarray):
    L, _ = layout()
    Ws, bs = [], []
    for wo, ws, bo, bl in L:
        Ws.append(params[wo:wo + ws[0] * ws[1]].reshape(ws))
        bs.append(params[bo:bo + bl])
    return Ws, bs


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even) through float32, returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def im2col(X: np.ndarray) -> np.ndarray:
    """X: (B, H, W, C) -> (B*Ho*Wo, 9*C) with columns ordered (ky, kx, c)."""
    B, H, W, Cc = X.shape
    Ho, Wo = conv_out(H), conv_out(W)
    Xp = np.zeros((B, H + 2, W + 2, Cc), dtype=X.dtype)
    Xp[:, 1:H + 1, 1:W + 1] = X
    cols = np.empty((B, Ho, Wo, 3, 3, Cc), dtype=X.dtype)
    for ky in range(3):
        for kx in range(3):
            cols[:, :, :, ky, kx, :] = Xp[:, ky:ky + 2 * Ho:2, kx:kx + 2 * Wo:2, :]
    return cols.reshape(B * Ho * Wo, 9 * Cc)


def col2im(dcol: np.ndarray, B: int, H: int, W: int, Cc: int) -> np.ndarray:
    Ho, Wo = conv_out(H), conv_out(W)
    d = dcol.reshape(B, Ho, Wo, 3, 3, Cc)
    dXp = np.zeros((B, H + 2, W + 2, Cc))
    for ky in range(3):
        for kx in range(3):
            dXp[:, ky:ky + 2 * Ho:2, kx:kx + 2 * Wo:2, :] += d[:, :, :, ky, kx, :]
    return dXp[:, 1:H + 1, 1:W + 1]


def forward(params, frames, pairs, prev, bf=True):
    """Returns (out (B,6), cache)."""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    Ws, bs = unpack(np.asarray(params, np.float64))
    frames = np.asarray(frames, np.float64)
    pairs = np.asarray(pairs)
    B = pairs.shape[0]
    X = np.stack([frames[pairs[:, 0]], frames[pairs[:, 1]]], -1)  # (B, H, W, 2)
    cache = {"cols": [], "Y": [], "X": [X]}
    for li, (cin, cout) in enumerate(CONVS):
        col = r(im2col(X))
        Y = col @ r(Ws[li]).T + bs[li]
        Y = np.maximum(Y, 0.0)
        Ho, Wo = conv_out(X.shape[1]), conv_out(X.shape[2])
        X = Y.reshape(B, Ho, Wo, cout)
        cache["cols"].append(col)
        cache["Y"].append(X)
    feat = X.reshape(B, -1, FEAT).mean(1)
    X5 = np.concatenate([feat, np.asarray(prev, np.float64)], 1)
    X5r = r(X5)
    H = np.maximum(X5r @ r(Ws[4]).T + bs[4], 0.0)
    Hr = r(H)
    O = Hr @ r(Ws[5]).T + bs[5] + np.asarray(prev, np.float64)  # skip connection
    cache.update(X5r=X5r, H=H, Hr=Hr, B=B)
    return O, cache


def gradients(params, frames, pairs, prev, target, bf=True):
    """Returns (loss, grad blob) of mean((out - target)^2)."""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    params = np.asarray(params, np.float64)
    Ws, _ = unpack(params)
    O, c = forward(params, frames, pairs, prev, bf)
    B = c["B"]
    T = np.asarray(target, np.float64)
    n = B * TWIST
    loss = float(((O - T) ** 2).sum() / n)
    dO = 2.0 * (O - T) / n
    dOr = r(dO)
    L, total = layout()
    g = np.zeros(total)

    def put(li, dW, db):
        wo, ws, bo, bl = L[li]
        g[wo:wo + ws[0] * ws[1]] = dW.reshape(-1)
        g[bo:bo + bl] = db

    put(5, dOr.T @ c["Hr"], dO.sum(0))
    dH = (dOr @ r(Ws[5])) * (c["H"] > 0)
    dHr = r(dH)
    put(4, dHr.T @ c["X5r"], dH.sum(0))
    dX5 = dHr @ r(Ws[4])[:, :FEAT]
    Y4 = c["Y"][3]
    dY = np.repeat(dX5[:, None, :], Y4.shape[1] * Y4.shape[2], 1).reshape(Y4.shape) / (Y4.shape[1] * Y4.shape[2])
    dY = dY * (Y4 > 0)
    for li in range(len(CONVS) - 1, -1, -1):
        cin, cout = CONVS[li]
        dYm = dY.reshape(-1, cout)
        dYr = r(dYm)
        put(li, dYr.T @ c["cols"][li], dYm.sum(0))
        if li == 0:
            break
        dcol = dYr @ r(Ws[li])
        Xin = c["Y"][li - 1]
        dX = col2im(dcol, B, Xin.shape[1], Xin.shape[2], cin)
        dY = dX * (Xin > 0)
    return loss, g


def init_params(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    L, total = layout()
    p = np.zeros(total)
    for li, (wo, ws, bo, bl) in enumerate(L):
        a = np.sqrt(6.0 / ws[1]) * (0.1 if li == 5 else 1.0)
        p[wo:wo + ws[0] * ws[1]] = rng.uniform(-a, a, ws[0] * ws[1])
        p[bo:bo + bl] = rng.uniform(-0.05, 0.05, bl)
    return p
