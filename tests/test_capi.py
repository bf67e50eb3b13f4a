"""C ABI: the libraries load, export every declared symbol, struct layouts agree
with the C compiler, and argument errors return status codes (no GPU needed)."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _exports(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


def test_every_declared_symbol_is_exported(kf, synth):
    declared = kf.exported_symbols()
    assert "kf_integrate" in declared and "kf_synth_render" in declared
    have = _exports(kf.LIB_PATH) | _exports(synth.SYNTH_PATH)
    missing = [s for s in declared if s not in have]
    assert not missing, missing


def test_native_library_is_gfx950():
    out = subprocess.run(["strings", os.path.join(ROOT, "openfusion_amd", "libkinfu.so")], capture_output=True,
                         text=True, check=True).stdout
    assert "gfx950" in out


def test_struct_layouts_match_c(kf, synth):
    structs = {
        "kf_intrinsics": kf.Intrinsics, "kf_volume_desc": kf.VolumeDesc, "kf_posef": kf.PoseF,
        "kf_pose": kf.Pose, "kf_level_maps": kf.LevelMaps, "kf_frame_maps": kf.FrameMaps,
        "kf_preprocess_params": kf.PreprocessParams, "kf_integrate_params": kf.IntegrateParams,
        "kf_raycast_params": kf.RaycastParams, "kf_icp_params": kf.IcpParams, "kf_icp_stats": kf.IcpStats,
        "kf_pipeline_desc": kf.PipelineDesc, "kf_frame_stats": kf.FrameStats,
        "kf_primitive": synth.Primitive, "kf_scene": synth.Scene, "kf_scene_desc": synth.SceneDesc,
        "kf_trajectory_desc": synth.TrajectoryDesc,
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "kinfu.h"', '#include "kf_synth.h"',
             "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("  return 0; }")
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "l.c")
        exe = os.path.join(td, "l")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-o", exe, src], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    got = dict(ln.rsplit(" ", 1) for ln in out.strip().splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(py, fname).offset, f"{cname}.{fname}"


def test_invalid_arguments_return_status(kf):
    L = kf.lib()
    h = C.c_void_p()
    assert L.kf_volume_create(None, C.byref(h)) == 1
    d = kf.VolumeDesc()
    d.n, d.extent_m, d.trunc_m, d.max_weight, d.format = 12, 1.0, 0.01, 128.0, 0  # n not a multiple of 8
    assert L.kf_volume_create(C.byref(d), C.byref(h)) == 1
    assert b"n must be" in L.kf_last_error()
    d.n, d.format = 64, 7
    assert L.kf_volume_create(C.byref(d), C.byref(h)) == 1
    k = kf.Intrinsics(525, 525, 319.5, 239.5, 8, 8)  # too small
    p = kf.IcpParams()
    assert L.kf_tracker_create(C.byref(k), C.byref(p), C.byref(h)) == 1
    k2 = kf.Intrinsics(525, 525, 319.5, 239.5, 640, 480)
    p.levels = 4
    assert L.kf_tracker_create(C.byref(k2), C.byref(p), C.byref(h)) == 1
    m = kf.FrameMaps()
    assert L.kf_maps_alloc(C.byref(k2), 0, 1, C.byref(m)) == 1
    assert L.kf_pipeline_process(None, None) == 1


def test_host_pose_fields(kf):
    from tests.helpers import pose
    T = pose(0.3, -0.2, 0.1, (0.5, -1.0, 2.0))
    p = kf.pose_fields_host(T)
    R = T[:3, :3]
    assert np.allclose(np.array(p.wc.r[:]).reshape(3, 3), R, atol=1e-7)
    assert np.allclose(np.array(p.cw.r[:]).reshape(3, 3), R.T, atol=1e-7)
    assert np.allclose(np.array(p.cw.t[:]), -R.T @ T[:3, 3], atol=1e-6)
    assert np.array_equal(np.array(p.wc.r[:], dtype=np.float32), R.reshape(-1).astype(np.float32))


def test_default_desc_matches_configuration(kf):
    d = kf.default_pipeline_desc()
    assert (d.intr.fx, d.intr.fy, d.intr.cx, d.intr.cy, d.intr.width, d.intr.height) == (525, 525, 319.5, 239.5,
                                                                                            640, 480)
    assert d.icp.levels == 3 and list(d.icp.iterations) == [10, 5, 4]
    assert list(d.icp_prior.iterations) == [4, 3, 3]
    assert abs(d.icp.dist_thresh_m - 0.1) < 1e-7 and d.icp.normal_thresh_deg == 20.0
    assert d.volume.max_weight == 128.0
    assert abs(d.fallback_residual_ratio - 1.5) < 1e-7 and abs(d.fallback_min_inlier_frac - 0.4) < 1e-7
