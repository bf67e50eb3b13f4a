"""Marching-tetrahedra oracle pinned by geometry: vertices on an analytic sphere,
area close to 4 pi r^2, a closed (edge-manifold) surface, outward orientation."""
import numpy as np

from oracle import mesh_oracle as mo


def _sphere(n=24, r=0.3, extent=1.0):
    h = extent / n
    c = (np.arange(n) + 0.5) * h - extent / 2
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    sd = np.sqrt(X * X + Y * Y + Z * Z) - r
    return np.clip(sd / 0.05, -1, 1), np.ones_like(sd), h


def test_sphere_mesh_geometry():
    f, w, h = _sphere()
    T = mo.marching_tetrahedra(f, w, h, (-0.5, -0.5, -0.5))
    assert T.shape[0] > 500
    rad = np.linalg.norm(T.reshape(-1, 3), axis=1)
    assert np.max(np.abs(rad - 0.3)) < 0.2 * h  # linear interpolation of a curved field
    area = 0.5 * np.linalg.norm(np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0]), axis=1).sum()
    assert abs(area - 4 * np.pi * 0.09) / (4 * np.pi * 0.09) < 0.03
    nrm = np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0])
    assert np.all((nrm * T.mean(1)).sum(1) > 0)  # faces the positive (outside) side


def test_sphere_mesh_is_closed():
    f, w, h = _sphere(16)
    T = mo.marching_tetrahedra(f, w, h, (-0.5, -0.5, -0.5))
    keys = np.round(T.reshape(-1, 3) / 1e-7).astype(np.int64)
    _, inv = np.unique(keys, axis=0, return_inverse=True)
    tri = inv.reshape(-1, 3)
    edges = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), 1)
    _, counts = np.unique(edges, axis=0, return_counts=True)
    degenerate = (tri[:, 0] == tri[:, 1]) | (tri[:, 1] == tri[:, 2]) | (tri[:, 0] == tri[:, 2])
    if not degenerate.any():
        assert np.all(counts == 2)


def test_unobserved_cells_skipped():
    f, w, h = _sphere(16)
    w[:8] = 0
    T = mo.marching_tetrahedra(f, w, h, (-0.5, -0.5, -0.5))
    assert np.all(T[..., 0] >= -0.5 + 7.5 * h - 1e-12)
