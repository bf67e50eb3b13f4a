"""Brick occupancy map of the volume (bricks of 8^3 voxels, each extended by one
voxel along +x, +y, +z): exact after a rebuild from the voxels, a superset of the
exact map while the sweep maintains it, and invisible in raycast output, which the
parity tests compare bit for bit with the oracle's full march."""
import numpy as np
import pytest
from scipy import ndimage

from tests.helpers import fill_volume, integ_params, vol_desc, voxel_centres

@pytest.fixture(scope="module")
def small(synth):
    return synth.make_stream("small", frames=5)


def exact_occupancy(f: np.ndarray) -> np.ndarray:
    """(nb, nb, nb) bool: some voxel of brick b's extended extent [8b, 8b + 8]^3
    (clipped to the volume) has f != 1."""
    a = f != np.float32(1.0)
    nb = a.shape[0] // 8
    for ax in range(3):
        a = np.moveaxis(a, ax, 0)
        core = a.reshape(nb, 8, *a.shape[1:]).any(axis=1)
        nxt = np.zeros_like(core)
        nxt[:-1] = a[8::8][: nb - 1]
        a = np.moveaxis(core | nxt, 0, ax)
    return a


def brick_distance_of(occ: np.ndarray, cap: int = 16) -> np.ndarray:
    """min(Chebyshev distance in bricks to the nearest set brick, cap) as uint8."""
    if not occ.any():
        return np.full(occ.shape, cap, np.uint8)
    d = ndimage.distance_transform_cdt(~occ, metric="chessboard")
    return np.minimum(d, cap).astype(np.uint8)


def _tsdf(kf, vol) -> np.ndarray:
    raw = vol.download_raw()
    n = vol.desc.n
    if vol.desc.format == kf.KF_VOXEL_F16:
        return (raw.view(np.uint32) & 0xFFFF).astype(np.uint16).view(np.float16).astype(np.float32).reshape(n, n, n)
    return raw.view(np.float32)[0::2].reshape(n, n, n)


def test_exact_occupancy_reference():
    # one voxel with f != 1 at (8, 3, 15): bricks (1|0, 0, 1)
    f = np.ones((32, 32, 32), np.float32)
    f[8, 3, 15] = 0.5
    occ = exact_occupancy(f)
    assert sorted(map(tuple, np.argwhere(occ))) == [(0, 0, 1), (1, 0, 1)]
    f[:] = 1
    f[8, 8, 8] = -1
    assert len(np.argwhere(exact_occupancy(f))) == 8


def test_brick_distance_reference_matches_brute_force():
    rng = np.random.default_rng(3)
    for nb, k in ((9, 4), (20, 2), (7, 0)):
        occ = np.zeros((nb,) * 3, bool)
        occ[tuple(rng.integers(0, nb, size=(3, k)))] = True
        g = np.stack(np.meshgrid(*(np.arange(nb),) * 3, indexing="ij"), -1).reshape(-1, 1, 3)
        pts = np.argwhere(occ).reshape(1, -1, 3)
        if k:
            brute = np.minimum(np.abs(g - pts).max(-1).min(-1), 16).reshape(occ.shape)
        else:
            brute = np.full(occ.shape, 16)
        assert np.array_equal(brick_distance_of(occ), brute)
    assert brick_distance_of(occ).max() == 16


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", [0, 1], ids=["f32", "f16"])
def test_brick_distance_matches_chessboard_transform(gpu, kf, fmt):
    """nb = 25: two 16-column tiles per axis, the second partial; scattered
    non-empty voxels so that distances reach the cap of 16."""
    n = 200
    d = vol_desc(n, 2.0, 0.05, fmt=fmt)
    f = np.ones((n, n, n), np.float32)
    for p in ((3, 5, 190), (130, 7, 9), (199, 199, 199), (64, 127, 100), (8, 8, 8)):
        f[p] = 0.25
    vol = kf.Volume(n, 2.0, 0.05, 128.0, fmt)
    vol.upload_raw(fill_volume(vol.desc, f, 1.0))
    occ = vol.occupancy()
    assert np.array_equal(occ, exact_occupancy(_tsdf(kf, vol)))
    dist = vol.brick_distance()
    ref = brick_distance_of(occ)
    assert dist.shape == (25, 25, 25) and ref.max() == 16 and ref.min() == 0
    assert np.array_equal(dist, ref)
    kf.General.Instance().InitVoxelArray(vol)
    kf.synchronize()
    assert np.all(vol.brick_distance() == 16)


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", [0, 1], ids=["f32", "f16"])
def test_rebuild_is_exact(gpu, kf, fmt):
    d = vol_desc(64, 1.0, 0.05, fmt=fmt, origin=(-0.5, -0.5, 0.4))
    P = voxel_centres(d)
    sd = np.linalg.norm(P - np.array([0.05, -0.02, 0.9]), axis=-1) - 0.27
    f = np.clip(sd / d.trunc_m, -1, 1).astype(np.float32)
    w = np.ones_like(f)
    w[:, :, :5] = 0
    vol = kf.Volume(64, 1.0, 0.05, 128.0, fmt, origin=(-0.5, -0.5, 0.4))
    vol.upload_raw(fill_volume(vol.desc, f, w))
    occ = vol.occupancy()
    ref = exact_occupancy(_tsdf(kf, vol))
    assert occ.shape == (8, 8, 8)
    assert 0 < ref.sum() < ref.size
    assert np.array_equal(occ, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", [0, 1], ids=["f32", "f16"])
def test_sweep_keeps_a_superset_and_reset_clears(gpu, kf, synth, small, fmt):
    k = synth.intrinsics()
    pr = synth.PRESETS["small"]
    vol = kf.Volume(pr.n, pr.extent, pr.trunc, pr.max_weight, fmt)
    assert not vol.occupancy().any() and np.all(vol.brick_distance() == 16)
    integ = kf.VolumeIntegrator(k, integ_params())
    for i in range(5):
        live = kf.DeviceMaps(k, 1, True)
        kf.CameraHandler.updateCameraData(k, synth.pipeline_desc("small").pre, live, kf.upload(small.depth[i]))
        integ.Integrate(vol, live, kf.DevicePose(small.poses[i]))
    kf.synchronize()
    occ = vol.occupancy()
    ref = exact_occupancy(_tsdf(kf, vol))
    assert ref.sum() > 100
    assert not (ref & ~occ).any()            # never misses a brick
    assert (occ & ~ref).sum() <= 0.01 * ref.sum()  # and rarely keeps a stale one
    assert np.array_equal(vol.brick_distance(), brick_distance_of(occ))
    vol.mark_modified()
    assert np.array_equal(vol.occupancy(), ref)
    kf.General.Instance().InitVoxelArray(vol)
    kf.synchronize()
    assert not vol.occupancy().any() and np.all(vol.brick_distance() == 16)
