"""End-to-end CPU run of the oracle pipeline on the small workload (256^3 FP32,
30 frames, arc of 2 cm / 1 deg per frame): it tracks, and the FP64 oracle agrees."""
import numpy as np
import pytest

from openfusion_amd import metrics


@pytest.fixture(scope="module")
def small(synth):
    return synth.make_stream("small")


def _run(orc, synth, stream, precision):
    d = synth.pipeline_desc("small")
    p = orc.Pipeline(d, precision)
    p.reset(stream.poses[0])
    est = [stream.poses[0]]
    for i in range(stream.depth.shape[0]):
        p.process(stream.depth[i])
        if i:
            est.append(p.pose())
    return np.array(est), p


def test_small_workload_tracks(orc, synth, small):
    est, p = _run(orc, synth, small, 0)
    ate = metrics.ate(est, small.poses)
    assert ate < 3e-3, ate
    st = p.stats()
    assert st.status == 0 and st.iterations == 19
    assert st.inliers > 0.4 * st.candidates


def test_fp64_oracle_agrees_with_fp32(orc, synth, small):
    frames = 10
    s = synth.Stream(small.preset, small.poses[:frames], small.scene, small.depth[:frames])
    e32, _ = _run(orc, synth, s, 0)
    e64, _ = _run(orc, synth, s, 1)
    for a, b in zip(e32, e64):
        assert metrics.translation_error(a, b) < 1e-3
        assert metrics.rotation_error_deg(a, b) < 0.05


def test_poses_stay_orthonormal(orc, synth, small):
    est, _ = _run(orc, synth, small, 0)
    for T in est:
        assert np.max(np.abs(T[:3, :3] @ T[:3, :3].T - np.eye(3))) < 1e-13


def test_rotation_error_metric_small_angles():
    from tests.helpers import pose
    for ang in (1e-7, 1e-5, 1e-3, 0.1, 1.0):
        a = pose(0.3, 0.2, 0.1)
        b = a @ pose(0.0, 0.0, ang)
        assert abs(metrics.rotation_error_deg(a, b) - np.degrees(ang)) < 1e-9 + 1e-9 * np.degrees(ang)
