"""Oracle ICP: identity -> zero increment, known rigid motion recovered, normal
equations symmetric PSD, gate counts vs FP64 brute force, Cholesky / Euler-ZYX /
constant-velocity algebra vs numpy."""
import math

import numpy as np

from openfusion_amd import metrics
from tests.helpers import icp_params, intr, pose, pre_params


def _scene_stream(synth, k, poses, seed=1000):
    sc = synth.scene((3.0, 3.0, 3.0), poses, n_spheres=3, n_boxes=2, seed=seed, clearance=0.4)
    return sc, [synth.render(sc, k, T, 5, i, noise_scale=0.0) for i, T in enumerate(poses)]


def _arc_pose(synth, i=0):
    """A pose of the small preset's arc: looks at the room centre, primitives in view."""
    return synth.trajectory(synth.KF_TRAJ_ARC, i + 1, 0.02, 1.0, (3, 3, 3), 0.7, 3000)[i]


def _normal_matrix(sums):
    A = np.zeros((6, 6))
    A[np.triu_indices(6)] = sums[:21]
    return A + np.triu(A, 1).T


def _model_from_live(orc, k, live, T, levels=None):
    """World-frame model maps = the live camera-frame maps moved by T."""
    levels = live.levels if levels is None else levels
    mod = orc.HostMaps(k, levels, False)
    R, t = T[:3, :3], T[:3, 3]
    for lvl in range(levels):
        v = live.vertex(lvl).reshape(-1, 3).astype(np.float64) @ R.T + t
        n = live.normal(lvl).reshape(-1, 3).astype(np.float64) @ R.T
        for i, nm in enumerate(("vx", "vy", "vz")):
            mod.set_plane(lvl, nm, v[:, i])
        for i, nm in enumerate(("nx", "ny", "nz")):
            mod.set_plane(lvl, nm, n[:, i])
    return mod


def test_identity_gives_zero_increment(orc, synth):
    k = intr()
    T = _arc_pose(synth)
    _, frames = _scene_stream(synth, k, np.stack([T]))
    live = orc.preprocess(k, pre_params(), frames[0], 3)
    model = _model_from_live(orc, k, live, T)
    sums, mask = orc.icp_reduce(k, 0, icp_params(), live, model, T, mask=True)
    valid = ~np.isnan(live.plane(0, "vz")) & ~np.isnan(live.plane(0, "nx"))
    assert sums[29] == valid.sum()  # every candidate ...
    assert sums[28] >= 0.999 * valid.sum()  # ... associates to itself (FP32 round trip aside)
    assert np.max(np.abs(sums[21:27])) < 1e-6 * sums[28]  # b ~ 0
    ev = np.linalg.eigvalsh(_normal_matrix(sums))
    assert ev.min() > 1e-6 * ev.max()  # the view constrains all six degrees of freedom
    ok, T2 = orc.solve_step(sums, T)
    assert ok
    assert metrics.translation_error(T, T2) < 1e-6 and metrics.rotation_error_deg(T, T2) < 1e-4


def test_normal_equations_symmetric_psd(orc, synth):
    k = intr()
    T0 = _arc_pose(synth)
    T1 = T0 @ pose(0.01, 0.02, -0.005, (0.015, 0.004, 0.01))
    _, fr = _scene_stream(synth, k, np.stack([T0, T1]))
    live = orc.preprocess(k, pre_params(), fr[1], 3)
    model = _model_from_live(orc, k, orc.preprocess(k, pre_params(), fr[0], 3), T0)
    sums, _ = orc.icp_reduce(k, 0, icp_params(), live, model, T0)
    ev = np.linalg.eigvalsh(_normal_matrix(sums))
    assert ev.min() > -1e-9 * ev.max()
    assert sums[28] > 10000


def test_known_rigid_motion_recovered(orc, synth):
    k = intr()
    T0 = _arc_pose(synth)
    dT = pose(math.radians(0.6), math.radians(-1.1), math.radians(0.4), (0.015, -0.008, 0.012))
    T1 = T0 @ dT
    _, fr = _scene_stream(synth, k, np.stack([T0, T1]))
    live = orc.preprocess(k, pre_params(), fr[1], 3)
    model = _model_from_live(orc, k, orc.preprocess(k, pre_params(), fr[0], 3), T0)
    est, st = orc.icp_track(k, icp_params(), live, model, T0)
    assert st.status == 0 and st.iterations == 19
    assert metrics.translation_error(est, T1) < 1e-3
    assert metrics.rotation_error_deg(est, T1) < 0.05


def test_gate_counts_vs_fp64_brute_force(orc, synth):
    k = intr(160, 120, fx=130.0, fy=130.0)
    T0 = _arc_pose(synth)
    T1 = T0 @ pose(0.02, 0.05, -0.01, (0.03, 0.01, 0.03))
    _, fr = _scene_stream(synth, k, np.stack([T0, T1]))
    p = icp_params(levels=1)
    live = orc.preprocess(k, pre_params(), fr[1], 1)
    model = _model_from_live(orc, k, orc.preprocess(k, pre_params(), fr[0], 1), T0)
    sums, mask = orc.icp_reduce(k, 0, p, live, model, T0, mask=True)
    # brute force in FP64
    s = model.vertex(0).reshape(-1, 3).astype(np.float64)
    m = model.normal(0).reshape(-1, 3).astype(np.float64)
    lz = live.plane(0, "vz").reshape(-1)
    cand = ~np.isnan(lz) & ~np.isnan(s[:, 0]) & ~np.isnan(m[:, 0])
    Ti = np.linalg.inv(T0)
    sc = s @ Ti[:3, :3].T + Ti[:3, 3]
    with np.errstate(all="ignore"):
        u = k.fx * sc[:, 0] / sc[:, 2] + k.cx
        v = k.fy * sc[:, 1] / sc[:, 2] + k.cy
    ru, rv = np.floor(u + 0.5), np.floor(v + 0.5)
    inb = cand & (sc[:, 2] > 1e-6) & (ru >= 0) & (ru < k.width) & (rv >= 0) & (rv < k.height)
    j = np.where(inb, rv * k.width + ru, 0).astype(np.int64)
    lv = live.vertex(0).reshape(-1, 3).astype(np.float64)[j]
    ln = live.normal(0).reshape(-1, 3).astype(np.float64)[j]
    ok = inb & ~np.isnan(lv[:, 0]) & ~np.isnan(ln[:, 0])
    dw = lv @ T0[:3, :3].T + T0[:3, 3]
    nw = ln @ T0[:3, :3].T
    dist = np.linalg.norm(dw - s, axis=1)
    cs = (nw * m).sum(1)
    thr = math.cos(math.radians(20.0))
    inl = ok & (dist < 0.1) & (cs > thr)
    edge = np.minimum.reduce([np.abs(u - np.floor(u) - 0.5), np.abs(v - np.floor(v) - 0.5),
                              np.abs(dist - 0.1) * 100, np.abs(cs - thr) * 100])
    amb = np.nan_to_num(edge, nan=1.0) < 1e-4
    got = mask.reshape(-1).astype(bool)
    assert sums[29] == cand.sum()
    assert np.array_equal(got[~amb], inl[~amb])
    assert abs(int(sums[28]) - int(inl.sum())) <= amb.sum()


def test_cholesky_euler_step_vs_numpy(orc):
    rng = np.random.default_rng(7)
    for _ in range(20):
        J = rng.normal(size=(200, 6)) * np.array([1, 1, 1, 0.5, 0.5, 0.5])
        b = rng.normal(size=200) * 1e-3
        A = J.T @ J
        g = J.T @ b
        sums = np.zeros(30)
        sums[:21] = A[np.triu_indices(6)]
        sums[21:27] = g
        sums[28] = 200
        T = pose(0.1, -0.2, 0.3, (0.5, -0.1, 0.2))
        ok, T2 = orc.solve_step(sums, T)
        assert ok
        x = np.linalg.solve(A, g)
        a, bb, gg = -x[0], -x[1], -x[2]
        Rz = np.array([[math.cos(gg), -math.sin(gg), 0], [math.sin(gg), math.cos(gg), 0], [0, 0, 1]])
        Ry = np.array([[math.cos(bb), 0, math.sin(bb)], [0, 1, 0], [-math.sin(bb), 0, math.cos(bb)]])
        Rx = np.array([[1, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])
        D = np.eye(4)
        D[:3, :3] = Rz @ Ry @ Rx
        D[:3, 3] = x[3:]
        assert np.max(np.abs(T2 - D @ T)) < 1e-12
        R = T2[:3, :3]
        assert np.max(np.abs(R @ R.T - np.eye(3))) < 1e-12


def test_singular_system_keeps_pose(orc):
    sums = np.zeros(30)
    sums[28] = 1000
    sums[0] = 1.0  # rank 1
    T = pose(0.1, 0.0, 0.0, (1, 2, 3))
    ok, T2 = orc.solve_step(sums, T)
    assert not ok and np.array_equal(T2, T)
    sums2 = np.zeros(30)
    sums2[:21] = np.eye(6)[np.triu_indices(6)]
    sums2[28] = 5  # fewer than 6 inliers
    ok, _ = orc.solve_step(sums2, T)
    assert not ok


def test_constant_velocity(orc):
    T0 = pose(0.1, 0.2, -0.1, (0.1, 0.2, 0.3))
    T1 = pose(0.12, 0.21, -0.09, (0.12, 0.19, 0.31))
    G = orc.constant_velocity(T0, T1)
    assert np.max(np.abs(G - T1 @ (np.linalg.inv(T0) @ T1))) < 1e-12
