"""Synthetic stream generator: closed forms, noise statistics, motion limits."""
import math

import numpy as np

from openfusion_amd import metrics
from tests.helpers import intr


def _empty_scene(synth, room=(3.0, 3.0, 3.0)):
    return synth.scene(room, np.eye(4)[None], 0, 0, (0, 0), 1)


def test_room_wall_depth_is_exact(synth):
    k = intr()
    sc = _empty_scene(synth)
    mm, z = synth.render(sc, k, np.eye(4), 7, 0, noise_scale=0.0, true_depth=True)
    # Camera at the room centre looking along +z: every ray hits the wall z = 1.5.
    assert np.all(z == np.float32(1.5))
    assert np.all(mm == 1500)


def test_sphere_closed_form(synth):
    k = intr()
    sc = _empty_scene(synth)
    sc.count = 1
    p = sc.prim[0]
    p.kind = synth.KF_PRIM_SPHERE
    p.center[0], p.center[1], p.center[2] = 0.1, -0.05, 1.0
    p.radius = 0.3
    _, z = synth.render(sc, k, np.eye(4), 7, 0, noise_scale=0.0, true_depth=True)
    xs = np.arange(k.width) - k.cx
    ys = np.arange(k.height) - k.cy
    X, Y = np.meshgrid(xs / k.fx, ys / k.fy)
    d = np.stack([X, Y, np.ones_like(X)], -1)
    c = np.array([0.1, -0.05, 1.0])
    a = (d * d).sum(-1)
    b = (d @ c)
    disc = b * b - a * (c @ c - 0.09)
    hit = disc >= 0
    t = (b - np.sqrt(np.where(hit, disc, 0))) / a
    zz = np.where(hit, t, 1.5)
    assert np.max(np.abs(z - zz)) < 1e-6


def test_noise_statistics(synth):
    k = intr()
    sc = _empty_scene(synth)
    mm = synth.render(sc, k, np.eye(4), 2000, 3, noise_scale=1.0)
    err = mm.astype(np.float64) / 1000.0 - 1.5
    sigma = 0.0012 + 0.0019 * (1.5 - 0.4) ** 2
    expect = math.sqrt(sigma ** 2 + (0.001 ** 2) / 12.0)  # plus uniform quantisation
    assert abs(err.std() - expect) / expect < 0.02
    assert abs(err.mean()) < 5e-5


def test_render_deterministic_and_frame_keyed(synth):
    k = intr()
    sc = _empty_scene(synth)
    a = synth.render(sc, k, np.eye(4), 5, 1)
    b = synth.render(sc, k, np.eye(4), 5, 1)
    c = synth.render(sc, k, np.eye(4), 5, 2)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)


def test_max_range_cutoff(synth):
    k = intr()
    sc = _empty_scene(synth, room=(12.0, 12.0, 12.0))
    mm, z = synth.render(sc, k, np.eye(4), 5, 0, noise_scale=0.0, true_depth=True, max_range=4.0)
    assert np.all(z >= 6.0)
    assert np.all(mm == 0)


def _check_limits(P, lt, lr_deg, tol=1e-9):
    for i in range(len(P) - 1):
        assert metrics.translation_error(P[i], P[i + 1]) <= lt + tol
        assert metrics.rotation_error_deg(P[i], P[i + 1]) <= lr_deg + 1e-6
    for T in P:
        R = T[:3, :3]
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-12)
        assert abs(np.linalg.det(R) - 1.0) < 1e-12


def test_arc_trajectory_exact_steps(synth):
    P = synth.trajectory(synth.KF_TRAJ_ARC, 30, 0.02, 1.0, (3, 3, 3), 0.7, 3000)
    for i in range(29):
        assert abs(metrics.translation_error(P[i], P[i + 1]) - 0.02) < 1e-12
        assert abs(metrics.rotation_error_deg(P[i], P[i + 1]) - 1.0) < 1e-6
    _check_limits(P, 0.02, 1.0)


def test_spline_and_burst_respect_limits(synth):
    P = synth.trajectory(synth.KF_TRAJ_SPLINE, 300, 0.03, 2.0, (3, 3, 3), 0.7, 3001)
    _check_limits(P, 0.03, 2.0)
    steps = [metrics.translation_error(P[i], P[i + 1]) for i in range(299)]
    assert np.mean(steps) > 0.005  # the camera actually moves
    B = synth.trajectory(synth.KF_TRAJ_BURST, 400, 0.06, 5.0, (6, 3, 6), 1.0, 3004)
    _check_limits(B, 0.06, 5.0)
    rot = [metrics.rotation_error_deg(B[i], B[i + 1]) for i in range(399)]
    tr = [metrics.translation_error(B[i], B[i + 1]) for i in range(399)]
    assert max(max(rot) / 5.0, max(tr) / 0.06) > 0.9  # bursts reach the limit
    # camera stays inside the room minus the margin
    c = P[:, :3, 3]
    assert np.all(np.abs(c) <= 1.5 - 0.7 + 1e-9)


def test_scene_clearance(synth):
    P = synth.trajectory(synth.KF_TRAJ_SPLINE, 100, 0.03, 2.0, (3, 3, 3), 0.7, 11)
    sc = synth.scene((3, 3, 3), P, n_random=(6, 10), seed=12, clearance=0.45)
    assert 6 <= sc.count <= 10
    cams = P[:, :3, 3]
    for i in range(sc.count):
        q = sc.prim[i]
        c = np.array(q.center[:])
        if q.kind == synth.KF_PRIM_SPHERE:
            dist = np.linalg.norm(cams - c, axis=1) - q.radius
        else:
            h = np.array(q.half[:])
            dd = np.abs(cams - c) - h
            dist = np.linalg.norm(np.maximum(dd, 0), axis=1) + np.minimum(dd.max(1), 0)
        assert dist.min() >= 0.45 - 1e-12
        ext = q.radius if q.kind == synth.KF_PRIM_SPHERE else np.array(q.half[:])
        assert np.all(np.abs(c) + ext <= 1.5 + 1e-12)


def test_presets_render(synth):
    s = synth.make_stream("small")
    assert s.depth.shape == (30, 480, 640)
    assert s.scene.count == 5
    assert (s.depth > 0).mean() > 0.9


def test_fill_volume_layouts():
    from tests.helpers import fill_volume, vol_desc
    f = np.linspace(-1, 1, 8 ** 3, dtype=np.float32).reshape(8, 8, 8)
    w = np.arange(8 ** 3, dtype=np.float32).reshape(8, 8, 8) % 7
    raw32 = fill_volume(vol_desc(8, fmt=0), f, w).view(np.float32).reshape(-1, 2)
    assert np.array_equal(raw32[:, 0], f.reshape(-1)) and np.array_equal(raw32[:, 1], w.reshape(-1))
    raw16 = fill_volume(vol_desc(8, fmt=1), f, 3.0).view(np.uint32)
    assert np.array_equal((raw16 & 0xFFFF).astype(np.uint16).view(np.float16), f.reshape(-1).astype(np.float16))
    assert np.all(raw16 >> 16 == 3)
