"""Volume export, marching-tetrahedra meshing and shaded rendering on the GPU
against their numpy restatements."""
import numpy as np
import pytest

from oracle import mesh_oracle as mo
from tests.helpers import fill_volume, intr, ray_params, vol_desc, voxel_centres

pytestmark = pytest.mark.gpu


def _sphere_volume(kf, fmt, n=32, r=0.3):
    d = vol_desc(n, 1.0, 0.05, fmt=fmt, origin=(-0.5, -0.5, 0.4))
    P = voxel_centres(d)
    sd = np.linalg.norm(P - np.array([0.0, 0.0, 0.9]), axis=-1) - r
    f = np.clip(sd / d.trunc_m, -1, 1).astype(np.float32)
    w = np.ones_like(f)
    w[:, :, :3] = 0  # a few unobserved slices
    vol = kf.Volume(n, 1.0, 0.05, 128.0, fmt, origin=(-0.5, -0.5, 0.4))
    vol.upload_raw(fill_volume(vol.desc, f, w))
    if fmt == kf.KF_VOXEL_F16:
        f = f.astype(np.float16).astype(np.float32)
    return vol, f, w


def _canon(T):
    T = np.asarray(T, np.float64)
    key = np.round(T.reshape(len(T), -1), 4)
    order = np.lexsort(key.T[::-1])
    return T[order]


@pytest.mark.parametrize("fmt", [0, 1], ids=["f32", "f16"])
def test_mesh_matches_oracle(gpu, kf, fmt):
    print(f"mesh fmt={fmt}", flush=True)
    vol, f, w = _sphere_volume(kf, fmt)
    T = vol.extract_mesh(200000)
    h = 1.0 / 32
    R = mo.marching_tetrahedra(f, w, h, (-0.5, -0.5, 0.4))
    assert T.shape == R.shape and T.shape[0] > 500
    assert np.max(np.abs(_canon(T) - _canon(R))) < 1e-5


@pytest.mark.parametrize("fmt", [0, 1], ids=["f32", "f16"])
def test_tsdf_export(gpu, kf, fmt):
    vol, f, w = _sphere_volume(kf, fmt)
    out = vol.export_tsdf()
    ref = np.where(w > 0, f * np.float32(0.05), np.float32(0.05))
    assert np.array_equal(out, ref.astype(np.float32))


def test_render_shaded_matches_oracle(gpu, kf):
    from tests.helpers import pose
    k = intr(160, 120, fx=130.0, fy=130.0)
    vol, _, _ = _sphere_volume(kf, 0, n=64)
    T = pose(0.02, -0.03, 0.0, (0.01, 0.0, 0.0))
    model = kf.DeviceMaps(k, 1, False)
    dp = kf.DevicePose(T)
    kf.Raycaster(k, ray_params()).Raycast(vol, dp, model)
    kf.synchronize()
    nx, ny, nz = (model.plane(0, c) for c in ("nx", "ny", "nz"))
    assert np.isfinite(nx).mean() > 0.1
    for mode in (0, 1):
        got = kf.render_shaded(k, model, dp, mode)
        ref = mo.shade(nx, ny, nz, T, k.fx, k.fy, k.cx, k.cy, mode)
        diff = np.abs(got.view(np.uint8).astype(int) - ref.view(np.uint8).astype(int))
        assert diff.max() <= 1
        assert np.array_equal(got == 0, ref == 0)
