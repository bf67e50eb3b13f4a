"""Recorded-stream format: round trip and a hand-written reference-style file."""
import os
import tempfile

import numpy as np


def test_round_trip(synth):
    from openfusion_amd import recording
    s = synth.make_stream("small", frames=3)
    with tempfile.TemporaryDirectory() as td:
        info = recording.write(td, s.depth, synth.intrinsics(), name="unit", date="2026-01-01", camera_type="x")
        assert sorted(os.listdir(td)) == ["Parameters.txt", "recorded1", "recorded2", "recorded3"]
        assert os.path.getsize(os.path.join(td, "recorded2")) == 640 * 480 * 2
        back = recording.read_info(td)
        assert (back.width, back.height, back.frames) == (640, 480, 3)
        assert (back.fx, back.fy, back.cx, back.cy) == (525.0, 525.0, 319.5, 239.5)
        assert back.name == b"unit" and back.camera_type == b"x"
        for i in range(3):
            assert np.array_equal(recording.read_depth(td, i + 1, back), s.depth[i])


def test_reads_reference_style_parameters_with_negative_focal_lengths(synth):
    from openfusion_amd import recording
    with tempfile.TemporaryDirectory() as td:
        open(os.path.join(td, "Parameters.txt"), "w").write(
            "320\n240\n-285.17 -285.17\n160 120\n0\n300\n6000\n2\nrealTrial\n2013-01-01-10-00-00\nPrimesense\n")
        frame = (np.arange(320 * 240) % 5000).astype("<u2").reshape(240, 320)
        frame.tofile(os.path.join(td, "recorded1"))
        info = recording.read_info(td)
        assert (info.width, info.height, info.frames) == (320, 240, 2)
        assert info.min_depth_mm == 300.0 and info.max_depth_mm == 6000.0
        k = recording.intrinsics(info)
        assert abs(k.fx - 285.17) < 1e-4 and k.cx == 160.0
        assert np.array_equal(recording.read_depth(td, 1, info), frame)
        try:
            recording.read_depth(td, 2, info)  # missing file
            raise AssertionError("expected IOError")
        except IOError:
            pass
