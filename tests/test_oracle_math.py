"""Pins the oracle's scalar building blocks against independent references."""
import math

import numpy as np


def test_exp_neg_matches_libm(orc):
    rng = np.random.default_rng(0)
    xs = np.concatenate([np.linspace(0.0, 30.0, 3001), rng.uniform(0.0, 79.9, 3000)]).astype(np.float32)
    worst = 0.0
    for x in xs:
        got = orc.lib().kfo_exp_neg(float(x))
        ref = math.exp(-float(x))
        worst = max(worst, abs(got - ref) / ref)
    assert worst < 1e-6, worst


def test_exp_neg_special_values(orc):
    L = orc.lib()
    assert L.kfo_exp_neg(0.0) == 1.0
    assert L.kfo_exp_neg(80.0) == 0.0
    assert L.kfo_exp_neg(float("nan")) == 0.0
    assert L.kfo_exp_neg(float("inf")) == 0.0


def test_float_to_half_matches_numpy_rne(orc):
    rng = np.random.default_rng(1)
    vals = np.concatenate([
        rng.uniform(-1, 1, 20000), rng.uniform(-70000, 70000, 2000),
        rng.uniform(-1e-4, 1e-4, 5000), rng.uniform(-1e-7, 1e-7, 2000),
        [0.0, -0.0, 1.0, -1.0, 65504.0, 65520.0, 65519.99, 6.1035156e-05, 5.9604645e-08, 2.9802322e-08,
         2.98023259e-08, 0.5 + 2 ** -12, 0.5 + 3 * 2 ** -12, float("inf"), -float("inf")],
    ]).astype(np.float32)
    ref = vals.astype(np.float16).view(np.uint16)
    L = orc.lib()
    got = np.array([L.kfo_float_to_half(float(v)) for v in vals], dtype=np.uint16)
    assert np.array_equal(got, ref)


def test_half_to_float_all_patterns(orc):
    L = orc.lib()
    bits = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    ref = bits.view(np.float16).astype(np.float32)
    got = np.array([L.kfo_half_to_float(int(b)) for b in bits], dtype=np.float32)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), ref[~nan].view(np.uint32))


def test_intrinsics_levels_closed_form(orc, kf):
    from tests.helpers import intr
    k0 = intr()
    for lvl in range(3):
        kl = orc.lib().kfo_intrinsics_level(k0, lvl)
        s = 2 ** lvl
        assert kl.width == 640 // s and kl.height == 480 // s
        assert kl.fx == 525.0 / s and kl.fy == 525.0 / s
        # pixel-centre aligned: centre of the image maps to the centre
        assert abs(kl.cx - ((kl.width - 1) / 2)) < 1e-6 and abs(kl.cy - ((kl.height - 1) / 2)) < 1e-6
        kg = kf.intrinsics_level(k0, lvl)
        assert (kg.fx, kg.fy, kg.cx, kg.cy, kg.width, kg.height) == (kl.fx, kl.fy, kl.cx, kl.cy, kl.width,
                                                                      kl.height)
