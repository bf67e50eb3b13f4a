"""Precision of the FP16+u16 volume with compensated FP32 / FP64 ICP against the FP64
oracle (FP64 maps, FP64 volume, FP64 ICP), on the 512^3 / 4 m workload:
per-frame pose drift of the full pipelines, TSDF error and raycast depth RMSE of
volumes integrated at identical poses.  The numbers are printed, and written to
precision_report.json in the directory named by KF_REPORT_DIR when that is set."""
import json
import os

import numpy as np
import pytest

from openfusion_amd import metrics
from tests.helpers import integ_params, ray_params

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

FRAMES = 12


def test_fp16_pipeline_vs_fp64_oracle(gpu, kf, orc, synth):
    s = synth.make_stream("full16", frames=FRAMES)
    k = synth.intrinsics()
    d16 = synth.pipeline_desc("full16")
    print("GPU FP16 pipeline", flush=True)
    p = kf.Pipeline(d16)
    dd = kf.upload(s.depth)
    p.reset(s.poses[0])
    for i in range(FRAMES):
        p.process(dd.ptr + i * s.depth[0].nbytes)
    est16 = p.trajectory()
    print("FP64 oracle pipeline", flush=True)
    o64 = orc.Pipeline(d16, 1)
    o64.reset(s.poses[0])
    est64 = [s.poses[0]]
    for i in range(FRAMES):
        o64.process(s.depth[i])
        if i:
            est64.append(o64.pose())
    est64 = np.array(est64)
    dt = [metrics.translation_error(a, b) for a, b in zip(est16, est64)]
    dr = [metrics.rotation_error_deg(a, b) for a, b in zip(est16, est64)]

    # Identical poses (the FP64 oracle's): GPU FP16 volume vs oracle FP64 volume.
    print("volumes at identical poses", flush=True)
    vol = kf.Volume(512, 4.0, 0.02, 128.0, kf.KF_VOXEL_F16)
    v64 = orc.new_volume(vol.desc, orc.FMT_F64)
    integ = kf.VolumeIntegrator(k, integ_params())
    for i in range(FRAMES):
        live = kf.DeviceMaps(k, 1, True)
        kf.CameraHandler.updateCameraData(k, d16.pre, live, kf.upload(s.depth[i]))
        integ.Integrate(vol, live, kf.DevicePose(est64[i]))
        kf.synchronize()
        # the FP64 oracle integrates its own FP64 raylen of the same frame
        om = orc.preprocess(k, d16.pre, s.depth[i], 1, precision=1)
        orc.integrate(vol.desc, orc.FMT_F64, v64, k, integ_params(), om.raylen(), est64[i])
    f16, w16 = vol.tsdf_weight()
    fw64 = v64.view(np.float64).reshape(-1, 2)
    f64, w64 = fw64[:, 0].reshape(f16.shape), fw64[:, 1].reshape(f16.shape)
    both = (w16 > 0) & (w64 > 0)
    err = np.abs(f16[both].astype(np.float64) - f64[both])  # in units of T
    same_mask = float(((w16 > 0) == (w64 > 0)).mean())
    # Voxels whose centre projects to a different pixel in FP32 and FP64 next to a
    # depth discontinuity take a different depth sample: counted apart.
    tsdf_flips = int((err > 1e-3).sum())

    model16 = kf.DeviceMaps(k, 1, False)
    kf.Raycaster(k, ray_params()).Raycast(vol, kf.DevicePose(est64[-1]), model16)
    kf.synchronize()
    m64, _ = orc.raycast(vol.desc, orc.FMT_F64, v64, k, ray_params(), est64[-1], 1)
    Ti = np.linalg.inv(est64[-1])
    v16 = model16.vertex(0).reshape(-1, 3).astype(np.float64)
    v64v = m64.vertex(0).reshape(-1, 3).astype(np.float64)
    ok = np.isfinite(v16[:, 0]) & np.isfinite(v64v[:, 0])
    z16 = (v16[ok] @ Ti[:3, :3].T + Ti[:3, 3])[:, 2]
    z64 = (v64v[ok] @ Ti[:3, :3].T + Ti[:3, 3])[:, 2]
    dz = np.abs(z16 - z64)
    same = dz < 0.05  # both rays stop on the same surface; else a silhouette flip
    rmse = float(np.sqrt(np.mean(dz[same] ** 2)))
    flips = int((~same).sum())

    rep = {
        "workload": "512^3 FP16+u16 over 4 m, T 0.02 m, frames 0-%d of the seeded full stream" % (FRAMES - 1),
        "pose_drift_max_mm": 1000 * max(dt), "pose_drift_max_deg": max(dr),
        "pose_drift_per_frame_mm": [round(1000 * x, 4) for x in dt],
        "tsdf_abs_err_over_T": {"max": float(err.max()), "p999": float(np.quantile(err, 0.999)),
                                "mean": float(err.mean())},
        "tsdf_abs_err_max_m": float(err.max() * 0.02),
        "tsdf_voxels_compared": int(both.sum()), "tsdf_voxels_over_1e-3": tsdf_flips,
        "observed_mask_agreement": same_mask,
        "raycast_depth_rmse_mm": 1000 * rmse, "raycast_depth_max_mm": 1000 * float(dz[same].max()),
        "raycast_pixels_compared": int(ok.sum()), "raycast_silhouette_flips": flips,
        "ate_gpu_fp16_mm": 1000 * metrics.ate(est16, s.poses), "ate_fp64_oracle_mm": 1000 * metrics.ate(est64, s.poses),
    }
    print(json.dumps(rep), flush=True)
    out_dir = os.environ.get("KF_REPORT_DIR")
    if out_dir and os.path.isdir(out_dir):
        with open(os.path.join(out_dir, "precision_report.json"), "w") as f:
            json.dump(rep, f, indent=1)
    assert max(dt) < 1e-3 and max(dr) < 0.05
    assert rep["tsdf_abs_err_over_T"]["p999"] < 1e-3
    assert tsdf_flips <= 1e-4 * both.sum()
    assert rmse < 5e-4
    assert flips <= 1e-3 * ok.sum()
    assert same_mask > 0.999
