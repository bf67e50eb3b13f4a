"""Trajectory accuracy metrics.

ATE (absolute trajectory error) is the RMSE of camera-centre error against ground
truth (Sturm et al., IROS 2012).  ``ate`` compares positions directly (frame 0 is
set to ground truth, so no alignment is needed); ``ate_aligned`` first applies
the least-squares rigid alignment of Umeyama (PAMI 1991) without scale.
"""
from __future__ import annotations

import numpy as np


def centres(poses: np.ndarray) -> np.ndarray:
    return np.asarray(poses, dtype=np.float64)[:, :3, 3]


def ate(est: np.ndarray, gt: np.ndarray) -> float:
    d = centres(est) - centres(gt)
    return float(np.sqrt(np.mean(np.sum(d * d, axis=1))))


def umeyama_rigid(src: np.ndarray, dst: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """R, t minimising sum |R src_i + t - dst_i|^2."""
    ms, md = src.mean(0), dst.mean(0)
    S = (dst - md).T @ (src - ms) / len(src)
    U, _, Vt = np.linalg.svd(S)
    D = np.eye(3)
    if np.linalg.det(U) * np.linalg.det(Vt) < 0:
        D[2, 2] = -1.0
    R = U @ D @ Vt
    return R, md - R @ ms


def ate_aligned(est: np.ndarray, gt: np.ndarray) -> float:
    a, b = centres(est), centres(gt)
    R, t = umeyama_rigid(a, b)
    d = (a @ R.T + t) - b
    return float(np.sqrt(np.mean(np.sum(d * d, axis=1))))


def rotation_error_deg(Ta: np.ndarray, Tb: np.ndarray) -> float:
    """Angle of Ra^T Rb from atan2(|skew part|, (trace - 1) / 2), well conditioned
    for small angles (arccos of the trace is not)."""
    R = Ta[:3, :3].T @ Tb[:3, :3]
    s = 0.5 * np.linalg.norm([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]])
    c = (np.trace(R) - 1.0) / 2.0
    return float(np.degrees(np.arctan2(s, c)))


def translation_error(Ta: np.ndarray, Tb: np.ndarray) -> float:
    return float(np.linalg.norm(Ta[:3, 3] - Tb[:3, 3]))
