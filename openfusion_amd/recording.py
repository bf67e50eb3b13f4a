"""Recorded depth streams in the reference's format (ctypes over include/kf_recording.h):
<dir>/Parameters.txt + raw uint16 <dir>/recorded<N> frames, N counted from 1."""
from __future__ import annotations

import ctypes as C
import dataclasses
import os

import numpy as np

from . import kinfu as kf
from . import synth


class RecordingInfo(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("fx", C.c_float), ("fy", C.c_float),
                ("cx", C.c_float), ("cy", C.c_float), ("z_offset_mm", C.c_float), ("min_depth_mm", C.c_float),
                ("max_depth_mm", C.c_float), ("frames", C.c_int32), ("name", C.c_char * 128),
                ("date", C.c_char * 64), ("camera_type", C.c_char * 64)]


_ready = False


def _lib():
    global _ready
    L = synth.slib()
    if not _ready:
        P = C.POINTER(RecordingInfo)
        L.kf_recording_read_info.argtypes = [C.c_char_p, P]
        L.kf_recording_write_info.argtypes = [C.c_char_p, P]
        L.kf_recording_read_depth.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]
        L.kf_recording_write_depth.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]
        L.kf_recording_intrinsics.argtypes = [P, C.POINTER(kf.Intrinsics)]
        for f in (L.kf_recording_read_info, L.kf_recording_write_info, L.kf_recording_read_depth,
                  L.kf_recording_write_depth, L.kf_recording_intrinsics):
            f.restype = C.c_int32
        _ready = True
    return L


def read_info(path: str) -> RecordingInfo:
    info = RecordingInfo()
    r = _lib().kf_recording_read_info(path.encode(), C.byref(info))
    if r:
        raise IOError(f"cannot read {os.path.join(path, 'Parameters.txt')} (status {r})")
    return info


def intrinsics(info: RecordingInfo) -> kf.Intrinsics:
    k = kf.Intrinsics()
    if _lib().kf_recording_intrinsics(C.byref(info), C.byref(k)):
        raise ValueError("invalid focal lengths in the recording")
    return k


def read_depth(path: str, frame: int, info: RecordingInfo) -> np.ndarray:
    out = np.empty((info.height, info.width), dtype=np.uint16)
    r = _lib().kf_recording_read_depth(path.encode(), frame, info.width, info.height, out.ctypes.data)
    if r:
        raise IOError(f"cannot read frame {frame} of {path} (status {r})")
    return out


def write(path: str, depth: np.ndarray, k: kf.Intrinsics, min_depth_mm=300.0, max_depth_mm=4000.0,
          name="synthetic", date="", camera_type="kf_synth") -> RecordingInfo:
    """Writes frames (F, H, W) uint16 as a reference-format recording."""
    os.makedirs(path, exist_ok=True)
    F, H, W = depth.shape
    L = _lib()
    for i in range(F):
        a = np.ascontiguousarray(depth[i], dtype=np.uint16)
        if L.kf_recording_write_depth(path.encode(), i + 1, W, H, a.ctypes.data):
            raise IOError(f"cannot write frame {i + 1}")
    info = RecordingInfo(W, H, k.fx, k.fy, k.cx, k.cy, 0.0, min_depth_mm, max_depth_mm, F, name.encode(),
                         date.encode(), camera_type.encode())
    if L.kf_recording_write_info(path.encode(), C.byref(info)):
        raise IOError("cannot write Parameters.txt")
    return info
