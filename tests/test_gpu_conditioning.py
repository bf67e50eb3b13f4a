"""Camera-side depth conditioning on the GPU vs its numpy restatement."""
import numpy as np
import pytest

from oracle import conditioning_oracle as co
from tests.helpers import intr

pytestmark = pytest.mark.gpu


def test_median_bit_exact(gpu, kf, synth):
    s = synth.make_stream("small", frames=1)
    d = s.depth[0].copy()
    rng = np.random.default_rng(0)
    d[rng.random(d.shape) < 0.3] = 0  # holes
    d[100:130, 200:260] = 0
    got = kf.depth_median(d)
    assert np.array_equal(got, co.median5(d))
    assert np.array_equal(kf.depth_median(d, 7), co.median5(d, 7))


def test_undistort_matches(gpu, kf, synth):
    s = synth.make_stream("small", frames=1)
    d = s.depth[0].copy()
    d[50:60, 50:70] = 0
    k = intr()
    dist = (0.05, -0.02, 0.001, -0.0015)
    got = kf.depth_undistort(d, k, dist)
    ref = co.undistort(d, k.fx, k.fy, k.cx, k.cy, *dist)
    assert np.array_equal(got == 0, ref == 0)
    assert np.abs(got.astype(int) - ref.astype(int)).max() <= 1
    # zero distortion is the identity where all 4 neighbours are valid
    ident = kf.depth_undistort(d, k, (0, 0, 0, 0))
    inner = ident[:-1, :-1] > 0
    assert np.array_equal(ident[:-1, :-1][inner], d[:-1, :-1][inner])
