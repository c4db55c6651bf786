"""Source voxelizers restated on the CPU (test oracle; see oracle/__init__.py).

Restates upstream ``src/caustica/sources.py:105-175`` (plane and bowl CW sources)
and ``src/caustica/analytic/geometry.py:15-64`` (Fibonacci spherical-cap sampling).

Pinning: bowl index sets against the upstream fixtures bowl_cfg0/1/2.npz and the
3-D solve fixtures' src_indices (tests/test_oracle_golden.py); cap invariants
(point count, area sum, all points on the sphere of radius roc about the focus)
in tests/test_oracle_physics.py.
"""

from __future__ import annotations

import numpy as np

GOLDEN_ANGLE = np.pi * (3.0 - np.sqrt(5.0))


def spherical_cap_points(aperture_radius: float, roc: float, spacing: float):
    a, f = float(aperture_radius), float(roc)
    if not 0 < a < f:
        raise ValueError("need 0 < aperture_radius < roc")
    cos_tmax = np.sqrt(1.0 - (a / f) ** 2)
    cap_area = 2.0 * np.pi * f**2 * (1.0 - cos_tmax)
    n = max(int(np.ceil(cap_area / spacing**2)), 16)
    j = np.arange(n, dtype=np.float64)
    cos_t = 1.0 - (1.0 - cos_tmax) * (j + 0.5) / n
    sin_t = np.sqrt(np.clip(1.0 - cos_t**2, 0.0, None))
    phi = j * GOLDEN_ANGLE
    pts = np.column_stack((f * sin_t * np.cos(phi), f * sin_t * np.sin(phi), f * (1.0 - cos_t)))
    areas = np.full(n, cap_area / n)
    return pts, areas


def bowl_indices(dx: float, aperture_radius: float, roc: float, apex, spacing=None) -> np.ndarray:
    ds = dx / 2.0 if spacing is None else float(spacing)
    pts, _ = spherical_cap_points(aperture_radius, roc, ds)
    idx = np.round(pts / dx).astype(np.int64) + np.asarray(apex, np.int64)
    return np.unique(idx, axis=0)


def plane_indices(shape, axis: int, pos: int) -> np.ndarray:
    other = [np.arange(n) for i, n in enumerate(shape) if i != axis]
    if other:
        mesh = np.meshgrid(*other, indexing="ij")
        cols = [m.reshape(-1) for m in mesh]
        n_pts = cols[0].size
    else:
        cols, n_pts = [], 1
    idx = np.empty((n_pts, len(shape)), np.int64)
    it = iter(cols)
    for d in range(len(shape)):
        idx[:, d] = pos if d == axis else next(it)
    return idx
