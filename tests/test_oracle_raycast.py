"""Oracle raycast: exact on planar TSDFs (trilinear is exact for linear fields),
bounded on a sphere, invalid where nothing is hit."""
import numpy as np

from tests.helpers import fill_volume, intr, pixel_rays, ray_params, vol_desc, voxel_centres


def _plane_volume(d, nrm, c, T):
    P = voxel_centres(d)
    sd = (P @ nrm) - c  # signed distance, positive on the camera side when nrm faces the camera
    return fill_volume(d, np.clip(sd / d.trunc_m, -1, 1), 1.0)


def test_fronto_plane_depth_exact(orc):
    k = intr(160, 120, fx=130.0, fy=130.0)
    d = vol_desc(64, 2.0, 0.06, origin=(-1.0, -1.0, 0.2))
    Zp = 1.3
    vox = _plane_volume(d, np.array([0.0, 0.0, -1.0]), -Zp, np.eye(4))
    m, steps = orc.raycast(d, d.format, vox, k, ray_params(), np.eye(4), 1)
    v, n = m.vertex(0), m.normal(0)
    valid = ~np.isnan(v[..., 2])
    assert valid.mean() > 0.95
    assert np.max(np.abs(v[valid][:, 2] - Zp)) < 2e-5
    X, Y = pixel_rays(k)
    assert np.max(np.abs(v[valid][:, 0] - (X * Zp)[valid])) < 2e-5
    assert np.max(np.abs(n[valid] - np.array([0, 0, -1.0]))) < 1e-4
    assert steps > 0


def test_tilted_plane_depth_exact_from_moved_camera(orc):
    from tests.helpers import pose
    k = intr(160, 120, fx=130.0, fy=130.0)
    d = vol_desc(64, 2.4, 0.08, origin=(-1.2, -1.2, -0.6))
    T = pose(0.05, -0.08, 0.03, (0.05, -0.02, -1.2))
    nrm = np.array([0.2, 0.1, -1.0])
    nrm /= np.linalg.norm(nrm)
    c = -0.4
    vox = _plane_volume(d, nrm, c, T)
    m, _ = orc.raycast(d, d.format, vox, k, ray_params(), T, 1)
    v, n = m.vertex(0).reshape(-1, 3), m.normal(0).reshape(-1, 3)
    valid = ~np.isnan(v[:, 0])
    assert valid.mean() > 0.9
    assert np.max(np.abs(v[valid] @ nrm - c)) < 3e-5  # on the plane
    # and on the pixel ray
    o = T[:3, 3]
    X, Y = pixel_rays(k)
    dirs = np.stack([X, Y, np.ones_like(X)], -1).reshape(-1, 3) @ T[:3, :3].T
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    rel = v[valid] - o
    perp = rel - (rel * dirs[valid]).sum(1, keepdims=True) * dirs[valid]
    assert np.max(np.linalg.norm(perp, axis=1)) < 3e-5
    assert np.max(np.abs(n[valid] - nrm)) < 2e-4


def test_sphere_depth_within_interpolation_bound(orc):
    k = intr(120, 90, fx=100.0, fy=100.0)
    d = vol_desc(96, 1.2, 0.04, origin=(-0.6, -0.6, 0.4))
    C, r = np.array([0.0, 0.0, 1.0]), 0.35
    P = voxel_centres(d)
    sd = np.linalg.norm(P - C, axis=-1) - r
    vox = fill_volume(d, np.clip(sd / d.trunc_m, -1, 1), 1.0)
    m, _ = orc.raycast(d, d.format, vox, k, ray_params(), np.eye(4), 1)
    v = m.vertex(0).reshape(-1, 3)
    valid = ~np.isnan(v[:, 0])
    assert valid.sum() > 1000
    err = np.abs(np.linalg.norm(v[valid] - C, axis=1) - r)
    assert err.max() < 5e-4  # 0.5 mm at h = 12.5 mm, r = 0.35 m


def test_no_hit_is_invalid(orc):
    k = intr(64, 48, fx=50.0, fy=50.0)
    d = vol_desc(32, 1.0, 0.05, origin=(-0.5, -0.5, 1.0))
    vox = fill_volume(d, np.ones((32, 32, 32)), 0.0)  # empty space
    m, _ = orc.raycast(d, d.format, vox, k, ray_params(), np.eye(4), 1)
    assert np.all(np.isnan(m.vertex(0)))
    # looking away from the volume
    T = np.diag([-1.0, 1.0, -1.0, 1.0])
    m2, steps = orc.raycast(d, d.format, vox, k, ray_params(), T, 1)
    assert np.all(np.isnan(m2.vertex(0))) and steps == 0


def test_model_pyramid_subsamples_level0(orc):
    k = intr(160, 120, fx=130.0, fy=130.0)
    d = vol_desc(64, 2.0, 0.06, origin=(-1.0, -1.0, 0.2))
    vox = _plane_volume(d, np.array([0.0, 0.0, -1.0]), -1.3, np.eye(4))
    m, _ = orc.raycast(d, d.format, vox, k, ray_params(), np.eye(4), 3)
    v0 = m.vertex(0)
    for lvl in (1, 2):
        s = 2 ** lvl
        ref = v0[::s, ::s][: m.maps.level[lvl].height, : m.maps.level[lvl].width]
        assert np.array_equal(np.isnan(ref), np.isnan(m.vertex(lvl)))
        ok = ~np.isnan(ref)
        assert np.array_equal(ref[ok], m.vertex(lvl)[ok])
