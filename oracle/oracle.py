"""ctypes wrapper of the CPU oracle (oracle/libkforacle.so).  TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never
by the product.  See kf_oracle.cpp for what each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(_HERE))
from openfusion_amd import kinfu as kf  # noqa: E402  (struct layouts only)

LIB_PATH = os.path.join(_HERE, "libkforacle.so")
FMT_F64 = 2

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        D = C.POINTER(C.c_double)
        V = C.c_void_p
        L.kfo_exp_neg.restype = C.c_float
        L.kfo_exp_neg.argtypes = [C.c_float]
        L.kfo_float_to_half.restype = C.c_uint16
        L.kfo_float_to_half.argtypes = [C.c_float]
        L.kfo_half_to_float.restype = C.c_float
        L.kfo_half_to_float.argtypes = [C.c_uint16]
        L.kfo_intrinsics_level.restype = kf.Intrinsics
        L.kfo_intrinsics_level.argtypes = [C.POINTER(kf.Intrinsics), C.c_int32]
        L.kfo_num_threads.restype = C.c_int32
        L.kfo_preprocess.restype = None
        L.kfo_preprocess.argtypes = [C.POINTER(kf.Intrinsics), C.POINTER(kf.PreprocessParams), V, C.c_int32,
                                     C.c_int32, C.POINTER(kf.FrameMaps)]
        L.kfo_volume_reset.restype = None
        L.kfo_volume_reset.argtypes = [C.POINTER(kf.VolumeDesc), C.c_int32, V]
        L.kfo_integrate.restype = None
        L.kfo_integrate.argtypes = [C.POINTER(kf.VolumeDesc), C.c_int32, V, C.POINTER(kf.Intrinsics),
                                    C.POINTER(kf.IntegrateParams), V, D, C.POINTER(C.c_uint64), V]
        L.kfo_raycast.restype = None
        L.kfo_raycast.argtypes = [C.POINTER(kf.VolumeDesc), C.c_int32, V, C.POINTER(kf.Intrinsics),
                                  C.POINTER(kf.RaycastParams), D, C.c_int32, C.POINTER(kf.FrameMaps),
                                  C.POINTER(C.c_uint64)]
        L.kfo_icp_reduce.restype = None
        L.kfo_icp_reduce.argtypes = [C.POINTER(kf.Intrinsics), C.c_int32, C.POINTER(kf.IcpParams),
                                     C.POINTER(kf.FrameMaps), C.POINTER(kf.FrameMaps), D, D, V]
        L.kfo_icp_solve_step.restype = C.c_int32
        L.kfo_icp_solve_step.argtypes = [D, D]
        L.kfo_icp_track.restype = None
        L.kfo_icp_track.argtypes = [C.POINTER(kf.Intrinsics), C.POINTER(kf.IcpParams), C.POINTER(kf.FrameMaps),
                                    C.POINTER(kf.FrameMaps), D, D, C.POINTER(kf.IcpStats)]
        L.kfo_constant_velocity.restype = None
        L.kfo_constant_velocity.argtypes = [D, D, D]
        L.kfo_pipeline_create.restype = V
        L.kfo_pipeline_create.argtypes = [C.POINTER(kf.PipelineDesc), C.c_int32]
        L.kfo_pipeline_destroy.restype = None
        L.kfo_pipeline_destroy.argtypes = [V]
        L.kfo_pipeline_reset.restype = None
        L.kfo_pipeline_reset.argtypes = [V, D]
        L.kfo_pipeline_process.restype = None
        L.kfo_pipeline_process.argtypes = [V, V]
        L.kfo_pipeline_get_pose.restype = None
        L.kfo_pipeline_get_pose.argtypes = [V, D]
        L.kfo_pipeline_get_stats.restype = None
        L.kfo_pipeline_get_stats.argtypes = [V, C.POINTER(kf.IcpStats)]
        L.kfo_pipeline_stage_ms.restype = None
        L.kfo_pipeline_stage_ms.argtypes = [V, D]
        L.kfo_pipeline_volume.restype = V
        L.kfo_pipeline_volume.argtypes = [V, C.POINTER(C.c_size_t)]
        _lib = L
    return _lib


def _d(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class HostMaps:
    """kf_frame_maps backed by host numpy planes (float32)."""

    def __init__(self, k0: kf.Intrinsics, levels: int = 3, with_depth: bool = True):
        self.k0 = k0
        self.levels = levels
        self.maps = kf.FrameMaps()
        self.maps.levels = levels
        self._keep = []
        fp = C.POINTER(C.c_float)

        def arr(n, fill):
            a = np.full(n, fill, dtype=np.float32)
            self._keep.append(a)
            return a, a.ctypes.data_as(fp)

        P0 = k0.width * k0.height
        self.raylen_a = None
        if with_depth:
            self.raylen_a, self.maps.raylen = arr(P0, 0.0)
        self.planes = []
        for l in range(levels):
            kl = lib().kfo_intrinsics_level(C.byref(k0), l)
            lv = self.maps.level[l]
            lv.width, lv.height = kl.width, kl.height
            P = kl.width * kl.height
            d = {}
            if with_depth:
                d["depth"], lv.depth = arr(P, 0.0)
            for nm in ("vx", "vy", "vz", "nx", "ny", "nz"):
                d[nm], p = arr(P, np.nan)
                setattr(lv, nm, p)
            self.planes.append(d)

    def plane(self, level: int, name: str) -> np.ndarray:
        lv = self.maps.level[level]
        return self.planes[level][name].reshape(lv.height, lv.width)

    def raylen(self) -> np.ndarray:
        return self.raylen_a.reshape(self.k0.height, self.k0.width)

    def vertex(self, level=0):
        return np.stack([self.plane(level, n) for n in ("vx", "vy", "vz")], -1)

    def normal(self, level=0):
        return np.stack([self.plane(level, n) for n in ("nx", "ny", "nz")], -1)

    def set_plane(self, level: int, name: str, a: np.ndarray) -> None:
        self.planes[level][name][:] = np.asarray(a, np.float32).reshape(-1)

    def set_raylen(self, a: np.ndarray) -> None:
        self.raylen_a[:] = np.asarray(a, np.float32).reshape(-1)


def preprocess(k0, params, depth_u16: np.ndarray, levels: int = 3, precision: int = 0) -> HostMaps:
    m = HostMaps(k0, levels, True)
    d = np.ascontiguousarray(depth_u16, dtype=np.uint16)
    lib().kfo_preprocess(C.byref(k0), C.byref(params), d.ctypes.data, levels, precision, C.byref(m.maps))
    return m


def voxel_bytes(fmt: int) -> int:
    return {kf.KF_VOXEL_F32: 8, kf.KF_VOXEL_F16: 4, FMT_F64: 16}[fmt]


def new_volume(desc: kf.VolumeDesc, fmt: int | None = None) -> np.ndarray:
    fmt = desc.format if fmt is None else fmt
    n = desc.n
    vox = np.zeros(n * n * n * voxel_bytes(fmt), dtype=np.uint8)
    lib().kfo_volume_reset(C.byref(desc), fmt, vox.ctypes.data)
    return vox


def integrate(desc, fmt, vox: np.ndarray, k0, params, raylen: np.ndarray, T: np.ndarray, mask: bool = False):
    n = desc.n
    upd = C.c_uint64(0)
    m = np.zeros((n * n * n + 31) // 32, dtype=np.uint32) if mask else None
    rl = np.ascontiguousarray(raylen, dtype=np.float32)
    T = np.ascontiguousarray(T, dtype=np.float64).reshape(16)
    lib().kfo_integrate(C.byref(desc), fmt, vox.ctypes.data, C.byref(k0), C.byref(params), rl.ctypes.data, _d(T),
                        C.byref(upd), m.ctypes.data if m is not None else None)
    return int(upd.value), m


def raycast(desc, fmt, vox: np.ndarray, k0, params, T: np.ndarray, levels: int = 3):
    m = HostMaps(k0, levels, False)
    steps = C.c_uint64(0)
    T = np.ascontiguousarray(T, dtype=np.float64).reshape(16)
    lib().kfo_raycast(C.byref(desc), fmt, vox.ctypes.data, C.byref(k0), C.byref(params), _d(T), levels,
                      C.byref(m.maps), C.byref(steps))
    return m, int(steps.value)


def icp_reduce(k0, level, icp_params, live: HostMaps, model: HostMaps, T: np.ndarray, mask: bool = False):
    sums = np.zeros(30, dtype=np.float64)
    lv = live.maps.level[level]
    mk = np.zeros(lv.width * lv.height, dtype=np.uint8) if mask else None
    T = np.ascontiguousarray(T, dtype=np.float64).reshape(16)
    lib().kfo_icp_reduce(C.byref(k0), level, C.byref(icp_params), C.byref(live.maps), C.byref(model.maps), _d(T),
                         _d(sums), mk.ctypes.data if mk is not None else None)
    return sums, (mk.reshape(lv.height, lv.width) if mk is not None else None)


def icp_track(k0, icp_params, live: HostMaps, model: HostMaps, guess: np.ndarray):
    out = np.zeros(16, dtype=np.float64)
    st = kf.IcpStats()
    g = np.ascontiguousarray(guess, dtype=np.float64).reshape(16)
    lib().kfo_icp_track(C.byref(k0), C.byref(icp_params), C.byref(live.maps), C.byref(model.maps), _d(g), _d(out),
                        C.byref(st))
    return out.reshape(4, 4), st


def solve_step(sums: np.ndarray, T: np.ndarray) -> tuple[bool, np.ndarray]:
    T = np.ascontiguousarray(T, dtype=np.float64).reshape(16).copy()
    s = np.ascontiguousarray(sums, dtype=np.float64)
    ok = lib().kfo_icp_solve_step(_d(s), _d(T))
    return bool(ok), T.reshape(4, 4)


def constant_velocity(T2: np.ndarray, T1: np.ndarray) -> np.ndarray:
    G = np.zeros(16)
    a = np.ascontiguousarray(T2, dtype=np.float64).reshape(16)
    b = np.ascontiguousarray(T1, dtype=np.float64).reshape(16)
    lib().kfo_constant_velocity(_d(a), _d(b), _d(G))
    return G.reshape(4, 4)


class Pipeline:
    """CPU pipeline (precision 0 = FP32 restatement, 1 = FP64 oracle)."""

    def __init__(self, desc: kf.PipelineDesc, precision: int = 0):
        self.desc = desc
        self.h = lib().kfo_pipeline_create(C.byref(desc), precision)

    def __del__(self):
        if getattr(self, "h", None):
            lib().kfo_pipeline_destroy(self.h)
            self.h = None

    def reset(self, T0):
        T = np.ascontiguousarray(T0, dtype=np.float64).reshape(16)
        lib().kfo_pipeline_reset(self.h, _d(T))

    def process(self, depth_u16: np.ndarray):
        d = np.ascontiguousarray(depth_u16, dtype=np.uint16)
        lib().kfo_pipeline_process(self.h, d.ctypes.data)

    def pose(self) -> np.ndarray:
        T = np.zeros(16)
        lib().kfo_pipeline_get_pose(self.h, _d(T))
        return T.reshape(4, 4)

    def stats(self) -> kf.IcpStats:
        s = kf.IcpStats()
        lib().kfo_pipeline_get_stats(self.h, C.byref(s))
        return s

    def stage_ms(self) -> np.ndarray:
        ms = np.zeros(4)
        lib().kfo_pipeline_stage_ms(self.h, _d(ms))
        return ms

    def volume_raw(self) -> np.ndarray:
        nb = C.c_size_t(0)
        p = lib().kfo_pipeline_volume(self.h, C.byref(nb))
        return np.ctypeslib.as_array((C.c_uint8 * nb.value).from_address(p)).copy()


def num_threads() -> int:
    return int(lib().kfo_num_threads())
