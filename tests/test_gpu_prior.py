"""Learned prior on the GPU against its FP64 oracle (BF16 operand rounding
emulated): forward outputs, parameter gradients of every layer, and an overfit
check of the Adam training step."""
import numpy as np
import pytest

from oracle import prior_oracle as po

pytestmark = pytest.mark.gpu


def _data(kf, rng, B=4, F=5):
    frames = rng.uniform(0.5, 3.5, size=(F, po.H0, po.W0)).astype(np.float32)
    frames[rng.random(frames.shape) < 0.1] = 0.0
    pairs = rng.integers(0, F, size=(B, 2)).astype(np.int32)
    prev = rng.normal(0, 0.5, size=(B, 6)).astype(np.float32)
    target = rng.normal(0, 0.5, size=(B, 6)).astype(np.float32)
    dev = [kf.upload(a) for a in (frames, pairs, prev, target)]
    return (frames, pairs, prev, target), dev


def test_forward_and_gradients_match_oracle(gpu, kf):
    rng = np.random.default_rng(0)
    (frames, pairs, prev, target), (df, dp, dv, dt) = _data(kf, rng)
    B = pairs.shape[0]
    prior = kf.Prior(max_batch=8)
    params = po.init_params(5).astype(np.float32)
    prior.set_params(params)
    dout = kf.DeviceBuffer(B * 6 * 4)
    print("prior forward B=4", flush=True)
    prior.forward(df, dp, dv, B, dout)
    kf.synchronize()
    out = dout.download(np.float32).reshape(B, 6)
    ref, _ = po.forward(params.astype(np.float64), frames, pairs, prev, bf=True)
    scale = np.abs(ref).max()
    assert np.max(np.abs(out - ref)) < 2e-3 * scale, (out, ref)
    # Inference reads layers 2-4 straight from the activations; the training forward
    # materialises im2col.  Same outputs: with the inference output as the target,
    # the training forward's loss is exactly zero.
    dl0 = kf.DeviceBuffer(4)
    prior.gradients(df, dp, dv, dout, B, kf.DeviceBuffer(prior.nparams * 4), dl0)
    kf.synchronize()
    assert float(dl0.download(np.float32)[0]) == 0.0
    print("prior gradients B=4", flush=True)
    dg = kf.DeviceBuffer(prior.nparams * 4)
    dl = kf.DeviceBuffer(4)
    prior.gradients(df, dp, dv, dt, B, dg, dl)
    kf.synchronize()
    g = dg.download(np.float32).astype(np.float64)
    loss = float(dl.download(np.float32)[0])
    rloss, rg = po.gradients(params.astype(np.float64), frames, pairs, prev, target, bf=True)
    assert abs(loss - rloss) < 1e-3 * rloss
    L, _ = po.layout()
    for li, (wo, ws, bo, bl) in enumerate(L):
        for lo, hi, what in ((wo, wo + ws[0] * ws[1], "W"), (bo, bo + bl, "b")):
            a, b = g[lo:hi], rg[lo:hi]
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
            assert rel < 5e-3, (li, what, rel)


def test_training_step_overfits_a_batch(gpu, kf):
    rng = np.random.default_rng(1)
    (_, pairs, _, _), (df, dp, dv, dt) = _data(kf, rng, B=16, F=20)
    prior = kf.Prior(max_batch=16)
    prior.init(7)
    dl = kf.DeviceBuffer(4)
    losses = []
    print("prior overfit 300 steps", flush=True)
    for i in range(300):
        prior.train_step(df, dp, dv, dt, 16, 1e-3, dl)
        if i % 50 == 0 or i == 299:
            kf.synchronize()
            losses.append(float(dl.download(np.float32)[0]))
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.1 * losses[0], losses
